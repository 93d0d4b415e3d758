// Plain FP64 baseline: the same math as the mixed path written the way the
// reference's CUDA library is written (11_Dec_2022/GPU_sph/
// Important_Test_Simulations/Evrard_collapse/myCppSPHLibs_v2.h): one thread
// per particle, <<<ceil(N/256), 256>>>, a full j = 0..N-1 loop over global
// memory per thread, FP64 throughout, no LDS and no cell structure.  It is the
// denominator of the mixed-vs-FP64 speedup and the FP64 GPU reference at sizes
// the CPU oracle cannot reach.  Neighbour decisions use the same separately
// rounded r^2 as the oracle, so counts and sets match it exactly.
#include "sphx_internal.h"

namespace {

constexpr int PT = 256;

inline unsigned grid_rows(int64_t r0, int64_t r1) { return (unsigned)((r1 - r0 + PT - 1) / PT); }

// smoothing_h (myCppSPHLibs_v2.h:7-64) with the canonical update rule, or
// (RULE 2) its own rule of fixed steps of hstep h_prev (also LIBSX3:947-1006),
// or (RULE 4) the LIBV7 driver's steps halved on each jump across the band
// (LIBV7:255-328), or (RULE 5) h_smooth_fast's 0.3 % steps on the current h
// with a strict count (LIBSX:1133-1184); the same decisions as
// k_smoothing<RULE>.
template <int RULE>
__global__ __launch_bounds__(PT) void kp_smoothing(
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    const double *__restrict__ h_prev, const int32_t *__restrict__ k_prev,
    const double *__restrict__ divv, const double *__restrict__ dt_p, double dt, int32_t nngb,
    int32_t ndown, int32_t nup, int32_t max_iter, double hstep, double hmin, int64_t n, int64_t r0,
    int64_t r1,
    double *__restrict__ h_out, int32_t *__restrict__ k_out, int32_t *__restrict__ it_out,
    int64_t *counters) {
#pragma clang fp contract(off)
    int64_t i = r0 + (int64_t)blockIdx.x * PT + threadIdx.x;
    if (i >= r1) return;
    if (dt_p) dt = *dt_p;   // the step's dt in device memory (adaptive time step)
    double xi = x[i], yi = y[i], zi = z[i];
    double hp = h_prev[i];
    double s = 2.0 * hp;
    if (RULE != 5 && k_prev && divv) {
        double f = sphx_h_factor(nngb, k_prev[i]);
        double t1 = 0.5 * hp;
        double t2 = 1.0 + f;
        double t3 = t1 * t2;
        double t6;
        if (RULE == 4) {         // LIBV7:265-266: ((1/3 h) divV) dt
            double u = (1.0 / 3.0) * hp;
            u = u * divv[i];
            t6 = u * dt;
        } else if (RULE == 2) {  // LIBSX3:961: (1/3 (h divV)) dt
            double u = hp * divv[i];
            u = (1.0 / 3.0) * u;
            t6 = u * dt;
        } else {
            double u = hp * divv[i];
            u = u * dt;
            t6 = u / 3.0;
        }
        double sp = 2.0 * (t3 + t6);
        if (sp > 0.0 && sp < INFINITY) s = sp;
    }
    double lo = 0.0, hi = 0.0;
    int64_t klo = 0, khi = 0, k = 0;
    bool have_lo = false, have_hi = false, conv = false;
    int it = 0;
    if (RULE >= 2) {
        double coef = hstep;  // (coeff * 2); RULE 4 halves it on each jump across the band
        double smax = s;
        int64_t kpre = nup;
        for (;;) {
            k = 0;
            for (int64_t j = 0; j < n; ++j) {
                double r2 = sphx_r2(xi - x[j], yi - y[j], zi - z[j]);
                if (RULE == 5 ? sqrt(r2) < s : sqrt(r2) <= s) ++k;
            }
            if (RULE == 4 && k > nup) k = nup + 1;  // LIBV7 stops counting there
            const bool in_band = k >= ndown && k <= nup;
            // LIBSX:1226-1228: the floor on h after the eleventh step
            const bool floor_h = RULE == 5 && hmin > 0.0 && it + 1 > 11 && 0.5 * s < hmin;
            if (RULE == 4 && ((k < ndown && kpre > nup) || (k > nup && kpre < ndown)))
                coef = 0.5 * coef;
            if (RULE == 5) {
                const double hc = 0.5 * s;
                if (k < ndown) s = 2.0 * (hc + 0.003 * hc);
                if (k > nup) s = 2.0 * (hc - 0.003 * hc);
            } else {
                if (k < ndown) s = s + coef * hp;
                if (k > nup) s = s - coef * hp;
            }
            if (s > smax) smax = s;
            kpre = k;
            ++it;
            if (floor_h) {
                s = 2.0 * hmin;
                conv = true;
                break;
            }
            if (in_band) {  // before the pass limit, as the reference tests
                conv = true;
                break;
            }
            if (it > (RULE == 4 ? SPHX_H4_PASSES : (RULE == 5 ? SPHX_H5_PASSES : max_iter))) {
                if (RULE == 2) s = smax;
                break;
            }
        }
    }
    while (RULE == 0 && it < max_iter) {
        double s2 = s * s;
        k = 0;
        for (int64_t j = 0; j < n; ++j) {
            double r2 = sphx_r2(xi - x[j], yi - y[j], zi - z[j]);
            if (r2 <= s2) ++k;
        }
        ++it;
        if (k >= ndown && k <= nup) {
            conv = true;
            break;
        }
        if (k < ndown) {
            lo = s; klo = k; have_lo = true;
            if (k >= n) break;
        } else {
            hi = s; khi = k; have_hi = true;
        }
        double snew = s * sphx_h_factor(nngb, (int)(k > 2147483647 ? 2147483647 : k));
        if (have_lo && have_hi) {
            if (!(snew > lo && snew < hi)) snew = 0.5 * (lo + hi);
            if (!(snew > lo && snew < hi)) break;
        }
        s = snew;
    }
    if (!conv) {
        if (RULE == 0) {
            if (have_hi) { s = hi; k = khi; }
            else { s = lo; k = klo; }
        }
        atomicAdd((unsigned long long *)&counters[CNT_H_CAPPED], 1ull);
    }
    h_out[i - r0] = 0.5 * s;
    k_out[i - r0] = (int32_t)k;
    if (it_out) it_out[i - r0] = it;
    atomicMax((unsigned long long *)&counters[CNT_H_ITERS_MAX], (unsigned long long)it);
    atomicAdd((unsigned long long *)&counters[CNT_H_ITERS_SUM], (unsigned long long)it);
}

__device__ __forceinline__ double kw(double q, double hij) {
    double h3 = hij * hij * hij;
    if (q <= 1.0) return SPHX_SIGMA_D / h3 * (1.0 - 1.5 * q * q + 0.75 * q * q * q);
    double w = fmax(2.0 - q, 0.0);
    return SPHX_SIGMA_D / h3 * 0.25 * w * w * w;
}

__device__ __forceinline__ double kf(double q, double hij) {
    double h5 = hij * hij * hij * hij * hij;
    if (q <= 1.0) return SPHX_SIGMA_D / h5 * (-3.0 + 2.25 * q);
    double w = fmax(2.0 - q, 0.0);
    return -0.75 * SPHX_SIGMA_D / h5 * w * w / q;
}

// getDensity (myCppSPHLibs_v2.h:69-114)
__global__ __launch_bounds__(PT) void kp_density(const double *__restrict__ x,
                                                 const double *__restrict__ y,
                                                 const double *__restrict__ z,
                                                 const double *__restrict__ m,
                                                 const double *__restrict__ h, int64_t n,
                                                 int64_t r0, int64_t r1, double *__restrict__ rho,
                                                 int32_t *__restrict__ nngb) {
    int64_t i = r0 + (int64_t)blockIdx.x * PT + threadIdx.x;
    if (i >= r1) return;
    double xi = x[i], yi = y[i], zi = z[i], hi = h[i];
    double s = 0.0;
    int32_t c = 0;
    for (int64_t j = 0; j < n; ++j) {
        double r2 = sphx_r2(xi - x[j], yi - y[j], zi - z[j]);
        double t = hi + h[j];
        if (r2 <= sphx_sq(t)) {
            double hij = 0.5 * t;
            double q = sqrt(r2) / hij;
            s += m[j] * kw(q, hij);
            ++c;
        }
    }
    rho[i - r0] = s;
    if (nngb) nngb[i - r0] = c;
}

// getPressure_Adiabatic / getCsound_Adiabatic (myCppSPHLibs_v2.h:766-789)
__global__ void kp_eos(const double *__restrict__ rho, const double *__restrict__ u, double gamma,
                       int64_t n, double *__restrict__ P, double *__restrict__ c) {
    int64_t i = (int64_t)blockIdx.x * PT + threadIdx.x;
    if (i >= n) return;
    P[i] = (gamma - 1.0) * rho[i] * u[i];
    c[i] = sqrt(gamma * (gamma - 1.0) * u[i]);
}

// div_curlVel (myCppSPHLibs_v2.h:247-322), signed divergence as in shear.py:8-69
__global__ __launch_bounds__(PT) void kp_divcurl(
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    const double *__restrict__ vx, const double *__restrict__ vy, const double *__restrict__ vz,
    const double *__restrict__ rho, const double *__restrict__ m, const double *__restrict__ h,
    int64_t n, int64_t r0, int64_t r1, double *__restrict__ divv, double *__restrict__ curlv) {
    int64_t i = r0 + (int64_t)blockIdx.x * PT + threadIdx.x;
    if (i >= r1) return;
    double xi = x[i], yi = y[i], zi = z[i], hi = h[i];
    double sd = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        double dx = xi - x[j], dy = yi - y[j], dz = zi - z[j];
        double r2 = sphx_r2(dx, dy, dz);
        double t = hi + h[j];
        if (r2 <= sphx_sq(t)) {
            double hij = 0.5 * t;
            double q = sqrt(r2) / hij;
            double F = kf(q, hij);
            double gx = F * dx, gy = F * dy, gz = F * dz;
            double vxij = vx[i] - vx[j], vyij = vy[i] - vy[j], vzij = vz[i] - vz[j];
            double w = m[j] / rho[i];
            sd += w * (-vxij * gx - vyij * gy - vzij * gz);
            cx += w * (vyij * gz - vzij * gy);
            cy += w * (vzij * gx - vxij * gz);
            cz += w * (vxij * gy - vyij * gx);
        }
    }
    divv[i - r0] = sd;
    curlv[i - r0] = sqrt(cx * cx + cy * cy + cz * cz);
}

struct PlainAV {
    double alpha, beta, eta, w_eps;
    int32_t av_form, balsara;
};

__device__ __forceinline__ double pbalsara(double dv, double cv, double c, double h) {
    double a = fabs(dv);
    return a / (a + cv + 0.0001 * c / h);
}

__device__ __forceinline__ double pav(const PlainAV &p, double vr, double r2, double hij,
                                      double ci, double cj, double rhoi, double rhoj, double fi,
                                      double fj) {
    if (!(vr <= 0.0)) return 0.0;
    double rhob = 0.5 * (rhoi + rhoj);
    double pi;
    if (p.av_form == 0) {
        double mu = hij * vr / (r2 + hij * hij * p.eta * p.eta);
        pi = (-p.alpha * 0.5 * (ci + cj) * mu + p.beta * mu * mu) / rhob;
    } else {
        double w = vr / (sqrt(r2) + p.w_eps);
        pi = -0.5 * p.alpha * (ci + cj - 3.0 * w) * w / rhob;
    }
    if (p.balsara) pi *= 0.5 * (fi + fj);
    return pi;
}

// acc_sph (myCppSPHLibs_v2.h:327-417)
__global__ __launch_bounds__(PT) void kp_acc_sph(Soa s, PlainAV av, int64_t n, int64_t r0,
                                                 int64_t r1, double *__restrict__ ax,
                                                 double *__restrict__ ay,
                                                 double *__restrict__ az) {
    int64_t i = r0 + (int64_t)blockIdx.x * PT + threadIdx.x;
    if (i >= r1) return;
    double xi = s.x[i], yi = s.y[i], zi = s.z[i], hi = s.h[i];
    double Pri = s.P[i] / (s.rho[i] * s.rho[i]);
    double fi = av.balsara ? pbalsara(s.divv[i], s.curlv[i], s.c[i], hi) : 1.0;
    double axt = 0.0, ayt = 0.0, azt = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        double dx = xi - s.x[j], dy = yi - s.y[j], dz = zi - s.z[j];
        double r2 = sphx_r2(dx, dy, dz);
        double t = hi + s.h[j];
        if (r2 <= sphx_sq(t)) {
            double hij = 0.5 * t;
            double q = sqrt(r2) / hij;
            double F = kf(q, hij);
            double vr = (s.vx[i] - s.vx[j]) * dx + (s.vy[i] - s.vy[j]) * dy + (s.vz[i] - s.vz[j]) * dz;
            double fj = av.balsara ? pbalsara(s.divv[j], s.curlv[j], s.c[j], s.h[j]) : 1.0;
            double pij = pav(av, vr, r2, hij, s.c[i], s.c[j], s.rho[i], s.rho[j], fi, fj);
            double Prj = s.P[j] / (s.rho[j] * s.rho[j]);
            double sc = s.m[j] * (Pri + Prj + pij) * F;
            axt -= sc * dx;
            ayt -= sc * dy;
            azt -= sc * dz;
        }
    }
    ax[i - r0] = axt;
    ay[i - r0] = ayt;
    az[i - r0] = azt;
}

// get_dU (myCppSPHLibs_v2.h:634-719)
__global__ __launch_bounds__(PT) void kp_du(Soa s, PlainAV av, int64_t n, int64_t r0, int64_t r1,
                                            double *__restrict__ du) {
    int64_t i = r0 + (int64_t)blockIdx.x * PT + threadIdx.x;
    if (i >= r1) return;
    double xi = s.x[i], yi = s.y[i], zi = s.z[i], hi = s.h[i];
    double Pri = s.P[i] / (s.rho[i] * s.rho[i]);
    double fi = av.balsara ? pbalsara(s.divv[i], s.curlv[i], s.c[i], hi) : 1.0;
    double dut = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        double dx = xi - s.x[j], dy = yi - s.y[j], dz = zi - s.z[j];
        double r2 = sphx_r2(dx, dy, dz);
        double t = hi + s.h[j];
        if (r2 <= sphx_sq(t)) {
            double hij = 0.5 * t;
            double q = sqrt(r2) / hij;
            double F = kf(q, hij);
            double vr = (s.vx[i] - s.vx[j]) * dx + (s.vy[i] - s.vy[j]) * dy + (s.vz[i] - s.vz[j]) * dz;
            double fj = av.balsara ? pbalsara(s.divv[j], s.curlv[j], s.c[j], s.h[j]) : 1.0;
            double pij = pav(av, vr, r2, hij, s.c[i], s.c[j], s.rho[i], s.rho[j], fi, fj);
            dut += s.m[j] * (Pri + 0.5 * pij) * F * vr;
        }
    }
    du[i - r0] = dut;
}

// acc_g (myCppSPHLibs_v2.h:119-168) with the FP64 formula of getAcc_g_smth
// (GPU_sph/IC/libsx2_2t.py:633-671) and the pair potential of getPE (593-629).
__global__ __launch_bounds__(PT) void kp_gravity(
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    const double *__restrict__ m, const double *__restrict__ eps, double G, int64_t n, int64_t r0,
    int64_t r1, double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az,
    double *__restrict__ pot) {
    int64_t i = r0 + (int64_t)blockIdx.x * PT + threadIdx.x;
    if (i >= r1) return;
    double xi = x[i], yi = y[i], zi = z[i], ei = eps[i];
    double accx = 0.0, accy = 0.0, accz = 0.0, ph = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        if (j == i) continue;
        double dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
        double rr = sqrt(dx * dx + dy * dy + dz * dz);
        double inv_r3 = 1.0 / (rr * rr * rr);
        double e = 0.5 * (ei + eps[j]);
        double q = rr / e;
        double fk, pk;
        if (q <= 1.0) {
            double q2 = q * q;
            fk = (1.0 / (e * e * e)) * ((4.0 / 3.0) - (6.0 / 5.0) * q2 + 0.5 * q2 * q);
            pk = (-2.0 / e) * ((1.0 / 3.0) * q2 - (3.0 / 20.0) * q2 * q2 + (1.0 / 20.0) * q2 * q2 * q) +
                 7.0 / 5.0 / e;
        } else if (q <= 2.0) {
            double q2 = q * q, q3 = q2 * q, q4 = q3 * q, q5 = q4 * q, q6 = q5 * q;
            fk = inv_r3 * ((-1.0 / 15.0) + (8.0 / 3.0) * q3 - 3.0 * q4 + (6.0 / 5.0) * q5 - (1.0 / 6.0) * q6);
            pk = (-1.0 / 15.0) / rr - (1.0 / e) * ((4.0 / 3.0) * q2 - q3 + 0.3 * q4 - (1.0 / 30.0) * q5) +
                 8.0 / 5.0 / e;
        } else {
            fk = inv_r3;
            pk = 1.0 / rr;
        }
        accx += G * fk * dx * m[j];
        accy += G * fk * dy * m[j];
        accz += G * fk * dz * m[j];
        ph -= G * m[j] * pk;
    }
    ax[i - r0] = accx;
    ay[i - r0] = accy;
    az[i - r0] = accz;
    if (pot) pot[i - r0] = ph;
}

// Neighbour lists under r_ij^2 <= (h_i + h_j)^2; pass 1 counts, pass 2 fills.
__global__ __launch_bounds__(PT) void kp_neighbors(const double *__restrict__ x,
                                                   const double *__restrict__ y,
                                                   const double *__restrict__ z,
                                                   const double *__restrict__ h, int64_t n,
                                                   int64_t r0, int64_t r1,
                                                   int64_t *__restrict__ offsets,
                                                   int32_t *__restrict__ idx) {
    int64_t i = r0 + (int64_t)blockIdx.x * PT + threadIdx.x;
    if (i >= r1) return;
    double xi = x[i], yi = y[i], zi = z[i], hi = h[i];
    int64_t o = idx ? offsets[i - r0] : 0;
    int64_t c = 0;
    for (int64_t j = 0; j < n; ++j) {
        double r2 = sphx_r2(xi - x[j], yi - y[j], zi - z[j]);
        double t = hi + h[j];
        if (r2 <= sphx_sq(t)) {
            if (idx) idx[o + c] = (int32_t)j;
            ++c;
        }
    }
    if (!idx) offsets[i - r0 + 1] = c;
}

}  // namespace

int plain_smoothing(const double *x, const double *y, const double *z, const double *h_prev,
                    const int32_t *k_prev, const double *divv, const double *dt_p, double dt,
                    const int32_t *, const sphx_params &p, int64_t n, int64_t r0, int64_t r1,
                    double *h_out, int32_t *k_out, int32_t *it_out, int64_t *counters,
                    hipStream_t st) {
    if (r1 <= r0) return SPHX_OK;
    const double hstep = sphx_h_step(p.h_mode);
    if (p.h_mode == 5)
        kp_smoothing<5><<<grid_rows(r0, r1), PT, 0, st>>>(x, y, z, h_prev, k_prev, divv, dt_p, dt,
                                                           p.nngb, p.ndown, p.nup, p.max_iter,
                                                           hstep, p.h_min, n, r0, r1, h_out, k_out,
                                                           it_out,
                                                           counters);
    else if (p.h_mode == 4)
        kp_smoothing<4><<<grid_rows(r0, r1), PT, 0, st>>>(x, y, z, h_prev, k_prev, divv, dt_p, dt,
                                                           p.nngb, p.ndown, p.nup, p.max_iter,
                                                           hstep, p.h_min, n, r0, r1, h_out, k_out,
                                                           it_out,
                                                           counters);
    else if (p.h_mode >= 2)
        kp_smoothing<2><<<grid_rows(r0, r1), PT, 0, st>>>(x, y, z, h_prev, k_prev, divv, dt_p, dt,
                                                           p.nngb, p.ndown, p.nup, p.max_iter,
                                                           hstep, p.h_min, n, r0, r1, h_out, k_out,
                                                           it_out,
                                                           counters);
    else
        kp_smoothing<0><<<grid_rows(r0, r1), PT, 0, st>>>(x, y, z, h_prev, k_prev, divv, dt_p, dt,
                                                           p.nngb, p.ndown, p.nup, p.max_iter,
                                                           hstep, p.h_min, n, r0, r1, h_out, k_out,
                                                           it_out,
                                                           counters);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int plain_density(const double *x, const double *y, const double *z, const double *m,
                  const double *h, int64_t n, int64_t r0, int64_t r1, double *rho,
                  int32_t *nngb, hipStream_t st) {
    if (r1 <= r0) return SPHX_OK;
    kp_density<<<grid_rows(r0, r1), PT, 0, st>>>(x, y, z, m, h, n, r0, r1, rho, nngb);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int plain_eos(const double *rho, const double *u, double gamma, int64_t n, double *P, double *c,
              hipStream_t st) {
    if (n <= 0) return SPHX_OK;
    kp_eos<<<grid_rows(0, n), PT, 0, st>>>(rho, u, gamma, n, P, c);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int plain_divcurl(const double *x, const double *y, const double *z, const double *vx,
                  const double *vy, const double *vz, const double *rho, const double *m,
                  const double *h, int64_t n, int64_t r0, int64_t r1, double *divv,
                  double *curlv, hipStream_t st) {
    if (r1 <= r0) return SPHX_OK;
    kp_divcurl<<<grid_rows(r0, r1), PT, 0, st>>>(x, y, z, vx, vy, vz, rho, m, h, n, r0, r1, divv,
                                                  curlv);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int plain_acc_sph(const Soa &s, const sphx_params &p, int64_t n, int64_t r0, int64_t r1,
                  double *ax, double *ay, double *az, hipStream_t st) {
    if (r1 <= r0) return SPHX_OK;
    PlainAV av{p.alpha, p.beta, p.eta, p.av_w_eps, p.av_form, p.balsara};
    kp_acc_sph<<<grid_rows(r0, r1), PT, 0, st>>>(s, av, n, r0, r1, ax, ay, az);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int plain_du(const Soa &s, const sphx_params &p, int64_t n, int64_t r0, int64_t r1, double *du,
             hipStream_t st) {
    if (r1 <= r0) return SPHX_OK;
    PlainAV av{p.alpha, p.beta, p.eta, p.av_w_eps, p.av_form, p.balsara};
    kp_du<<<grid_rows(r0, r1), PT, 0, st>>>(s, av, n, r0, r1, du);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int plain_gravity(const double *x, const double *y, const double *z, const double *m,
                  const double *eps, double G, int64_t n, int64_t r0, int64_t r1, double *ax,
                  double *ay, double *az, double *pot, hipStream_t st) {
    if (r1 <= r0) return SPHX_OK;
    kp_gravity<<<grid_rows(r0, r1), PT, 0, st>>>(x, y, z, m, eps, G, n, r0, r1, ax, ay, az, pot);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int plain_neighbors(const double *x, const double *y, const double *z, const double *h,
                    int64_t n, int64_t r0, int64_t r1, int64_t *offsets, int32_t *idx,
                    hipStream_t st) {
    if (r1 <= r0) return SPHX_OK;
    kp_neighbors<<<grid_rows(r0, r1), PT, 0, st>>>(x, y, z, h, n, r0, r1, offsets, idx);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}
