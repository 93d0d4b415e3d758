// FP64 integrator (KDK leapfrog + u trapezoid), the implicit per-particle
// cooling solve, energy/momentum diagnostics and small utility kernels.
//
// Reference: v_evolve / r_evolve / u_updater_main / u_ut_previous_updater
// (11_Dec_2022/GPU_sph/Important_Test_Simulations/Evrard_collapse/
// myCppSPHLibs_v2.h:441-486, 724-761) driven as in hfv_sph_gpu_v2_evrard.cu:
// 407-482; DoCooling_h (GPU_sph/Cooling_v1/photolibs3.py:489-543); getKE /
// getPE (GPU_sph/IC/libsx2_2t.py:579-629).
#include "sphx_internal.h"

#include <mutex>
#include <vector>

namespace {

constexpr int IT = 256;
inline unsigned g1(int64_t n) { return (unsigned)((n + IT - 1) / IT); }

#define MH_CGS 1.6726e-24
#define KB_CGS 1.3807e-16

__global__ void k_zero_i64(int64_t *p, int n) {
    int i = threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_kick_drift(Soa s, const RunParams *__restrict__ rp, int64_t n) {
#pragma clang fp contract(off)
    int64_t i = (int64_t)blockIdx.x * IT + threadIdx.x;
    if (i >= n) return;
    double dt = rp[s.run[i]].dt;
    double half = 0.5 * dt;
    double vx = s.vx[i] + s.ax[i] * half;
    double vy = s.vy[i] + s.ay[i] * half;
    double vz = s.vz[i] + s.az[i] * half;
    s.vx[i] = vx; s.vy[i] = vy; s.vz[i] = vz;
    s.x[i] = s.x[i] + vx * dt;
    s.y[i] = s.y[i] + vy * dt;
    s.z[i] = s.z[i] + vz * dt;
}

__global__ void k_combine(Soa s, int64_t p0, int64_t p1) {
#pragma clang fp contract(off)
    int64_t i = p0 + (int64_t)blockIdx.x * IT + threadIdx.x;
    if (i >= p1) return;
    s.ax[i] = s.gx[i] + s.ax[i];
    s.ay[i] = s.gy[i] + s.ay[i];
    s.az[i] = s.gz[i] + s.az[i];
}

// Lambda(T): linear in T inside a bin, clamped to the end values outside.
__device__ double cool_lambda(double T, const double *__restrict__ Tt,
                              const double *__restrict__ Lt, int nt) {
#pragma clang fp contract(off)
    if (T <= Tt[0]) return Lt[0];
    if (T >= Tt[nt - 1]) return Lt[nt - 1];
    int lo = 0, hi = nt - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (Tt[mid] <= T) lo = mid;
        else hi = mid;
    }
    double f = (T - Tt[lo]) / (Tt[hi] - Tt[lo]);
    return Lt[lo] + (Lt[hi] - Lt[lo]) * f;
}

__device__ double cool_g(double u, double u_old, double dt_s, double nH, double rho_cgs,
                         double tfac, const CoolDev &cd) {
#pragma clang fp contract(off)
    double T = tfac * u;
    double lam = cool_lambda(T, cd.T, cd.L, cd.ntab);
    double rate = (nH * cd.heat - nH * nH * lam) / rho_cgs;
    return u - u_old - dt_s * rate;
}

// Backward-Euler root of g(u) = 0: bracket by factors of 1.1 from u_old (the
// search of photolibs3.py:505-521), then bisection to a relative width tol.
// The gas never cools below the table's first temperature node: the input is
// floored there, the bracket stops there, and a net cooling rate at the floor
// returns the floor.
__device__ double cool_one(double rho, double u_code, double dt, const CoolDev &cd, int *iters) {
#pragma clang fp contract(off)
    double rho_cgs = rho * cd.unit_rho_cgs;
    double dt_s = dt * cd.unit_time_s;
    double nH = cd.XH * rho_cgs / MH_CGS;
    double tfac = (cd.gamma - 1.0) * cd.mu * MH_CGS / KB_CGS;
    double u_floor = cd.T[0] / tfac;
    double u_old = u_code * cd.unit_u_cgs;
    if (u_old < u_floor) u_old = u_floor;
    *iters = 0;
    if (!(dt_s > 0.0)) return u_old / cd.unit_u_cgs;
    double g0 = cool_g(u_old, u_old, dt_s, nH, rho_cgs, tfac, cd);
    if (g0 == 0.0) return u_old / cd.unit_u_cgs;
    double lo, hi;
    if (g0 > 0.0) {
        if (cool_g(u_floor, u_old, dt_s, nH, rho_cgs, tfac, cd) >= 0.0) return u_floor / cd.unit_u_cgs;
        hi = u_old;
        lo = u_old / 1.1;
        if (lo < u_floor) lo = u_floor;
        int nb = 0;
        while (cool_g(lo, u_old, dt_s, nH, rho_cgs, tfac, cd) > 0.0 && nb < cd.max_bracket) {
            hi = lo;
            lo = lo / 1.1;
            if (lo < u_floor) lo = u_floor;
            ++nb;
        }
    } else {
        lo = u_old;
        hi = u_old * 1.1;
        int nb = 0;
        while (cool_g(hi, u_old, dt_s, nH, rho_cgs, tfac, cd) < 0.0 && nb < cd.max_bracket) {
            lo = hi;
            hi = hi * 1.1;
            ++nb;
        }
    }
    int it = 0;
    while ((hi - lo) > cd.tol * lo && it < cd.max_iter) {
        double mid = 0.5 * (lo + hi);
        if (cool_g(mid, u_old, dt_s, nH, rho_cgs, tfac, cd) > 0.0) hi = mid;
        else lo = mid;
        ++it;
    }
    *iters = it;
    return 0.5 * (lo + hi) / cd.unit_u_cgs;
}

// ---------------------------------------------------------------------------
// Analytic H/He network (GPU_sph/Cooling_v1/photolibs3.py:140-543), FP64:
// Cen (1992) recombination and collisional rates, photoionisation equilibrium
// of H and He under the background rates nd.rf (a damped fixed point in the
// electron fraction), T(u) through the mean molecular weight (a damped fixed
// point), the net rate at T clamped to 1e4..1e8 K, and DoCooling_h's
// backward-Euler solve: bracket the root by factors of 1.1 from u_old, then
// bisect until the bracket is within tol of the midpoint.
// ---------------------------------------------------------------------------
#define NET_KB 1.3807e-16
#define NET_MH 1.6726e-24

struct NetCI {
    double aHp, aHep, aHepp, ad, geH0, geHe0, geHep;
};

__device__ __forceinline__ NetCI net_ci(double T) {
    NetCI r;
    double Tfact = 1.0 / (1.0 + sqrt(T / 1e5));
    r.aHp = 8.41e-11 * pow(T / 1000.0, -0.2) / (1.0 + pow(T / 1e6, 0.7)) / sqrt(T);
    r.aHep = 1.5e-10 * pow(T, -0.6353);
    r.aHepp = 4.0 * r.aHp;
    r.ad = 1.9e-3 * pow(T, -1.5) * exp(-470000.0 / T) * (1.0 + 0.3 * exp(-94000.0 / T));
    r.geH0 = 5.85e-11 * sqrt(T) * exp(-157809.1 / T) * Tfact;
    r.geHe0 = 2.38e-11 * sqrt(T) * exp(-285335.4 / T) * Tfact;
    r.geHep = 5.68e-12 * sqrt(T) * exp(-631515.0 / T) * Tfact;
    return r;
}

// Abundances relative to nH: nH0, nHe0, nHp, ne, nHep, nHepp (helium by mass
// fixed at Y = 0.24, as the reference).
__device__ __forceinline__ void net_abundance(const NetCI &r, double nH, const double *rf,
                                              double *ab) {
    const double y = 0.24 / (4.0 - 4.0 * 0.24);
    double ne = 1.0, ne_old;
    double nH0 = 0, nHp = 0, nHep = 0, nHe0 = 0, nHepp = 0;
    int niter = 1;
    double dne = 1.0;
    while (dne > 1e-4 && niter < 100) {
        ne_old = ne;
        double ph = ne * nH;
        nH0 = r.aHp / (r.aHp + r.geH0 + rf[0] / ph);
        nHp = 1.0 - nH0;
        double a = (r.aHep + r.ad) / (r.geHe0 + rf[1] / ph);
        double b = (r.geHep + rf[2] / ph) / r.aHepp;
        nHep = y / (1.0 + a + b);
        nHe0 = nHep * (r.aHep + r.ad) / (r.geHe0 + rf[1] / ph);
        nHepp = nHep * (r.geHep + rf[2] / ph) / r.aHepp;
        ne = nHp + nHep + 2.0 * nHepp;
        ne = 0.5 * (ne + ne_old);
        dne = fabs(ne - ne_old);
        ++niter;
    }
    ab[0] = nH0; ab[1] = nHe0; ab[2] = nHp; ab[3] = ne; ab[4] = nHep; ab[5] = nHepp;
}

__device__ double net_u_to_temp(double u, double nH, double XH, const double *rf) {
    const double gamma = 5.0 / 3.0;
    double yHe = (1.0 - XH) / (4.0 * XH);
    double mu = (1.0 + 4.0 * yHe) / (1.0 + yHe + 1.0);
    double temp = (gamma - 1.0) * NET_MH / NET_KB * mu * u;
    int niter = 1;
    double dT = temp;
    double ab[6];
    while (dT / temp > 1e-4 && niter < 100) {
        double temp_old = temp;
        net_abundance(net_ci(temp_old), nH, rf, ab);
        mu = (1.0 + 4.0 * yHe) / (1.0 + yHe + ab[3]);
        temp = (gamma - 1.0) * NET_MH / NET_KB * mu * u;
        temp = 0.5 * (temp + temp_old);
        dT = fabs(temp - temp_old);
        ++niter;
    }
    return temp;
}

// Heating minus cooling per nH^2 at the temperature of u (clamped to 1e4..1e8 K).
__device__ double net_rate(double u, double nH, double XH, const double *rf) {
    double T = net_u_to_temp(u, nH, XH, rf);
    T = T < 1e4 ? 1e4 : (T > 1e8 ? 1e8 : T);
    NetCI r = net_ci(T);
    double ab[6];
    net_abundance(r, nH, rf, ab);
    double nH0 = ab[0], nHe0 = ab[1], nHp = ab[2], ne = ab[3], nHep = ab[4], nHepp = ab[5];
    double heat = nH0 * rf[3] / nH + nHe0 * rf[4] / nH + nHep * rf[5] / nH;
    double Tfact = 1.0 / (1.0 + sqrt(T / 1e5));
    double BetaH0 = 7.5e-19 * exp(-118348.0 / T) * Tfact;
    double BetaHep = 5.54e-17 * pow(T, -0.397) * exp(-473638.0 / T) * Tfact;
    double lt = 5.5 - log10(T);
    double Betaff = 1.43e-27 * sqrt(T) * (1.0 + 0.34 * exp(-(lt * lt) / 3.0));
    double exc = BetaH0 * ne * nH0 + BetaHep * ne * nHep;
    double ion = 2.18e-11 * r.geH0 * ne * nH0 + 3.94e-11 * r.geHe0 * ne * nHe0 +
                 8.72e-11 * r.geHep * ne * nHep;
    double rec = 1.036e-16 * T * ne * (r.aHp * nHp) + 1.036e-16 * T * ne * (r.aHep * nHep) +
                 1.036e-16 * T * ne * (r.aHepp * nHepp) + 6.526e-11 * r.ad * ne * nHep;
    double ff = Betaff * (nHp + nHep + 4.0 * nHepp) * ne;
    return heat - (exc + ion + rec + ff);
}

// DoCooling_h in cgs; *iters = bisection iterations, *capped = 1 when a cap
// ended the bracket or the bisection.
__device__ double net_cool_one(double rho, double u_old, double dt, const NetDev &nd, int *iters,
                               int *capped) {
    const double nH = nd.XH * rho / NET_MH;
    const double ratefact = nH * nH / rho;
    double u = u_old, u_upper = u, u_lower = u;
    *iters = 0;
    *capped = 0;
    double g0 = u - u_old - ratefact * net_rate(u, nH, nd.XH, nd.rf) * dt;
    if (g0 == 0.0) return u_old;
    int nb = 0;
    if (g0 > 0.0) {
        while (u_upper - u_old - ratefact * net_rate(u_upper, nH, nd.XH, nd.rf) * dt > 0.0) {
            if (nb++ >= nd.max_bracket) { *capped = 1; break; }
            u_upper /= 1.1;
        }
        u_lower = u_upper;
        u_upper = u_lower * 1.1;
    } else {
        while (u_upper - u_old - ratefact * net_rate(u_upper, nH, nd.XH, nd.rf) * dt < 0.0) {
            if (nb++ >= nd.max_bracket) { *capped = 1; break; }
            u_upper *= 1.1;
        }
        u_lower = u_upper / 1.1;
    }
    int niter = 1;
    double du = u;
    while (fabs(du / u) > nd.tol && niter < nd.max_iter) {
        u = 0.5 * (u_lower + u_upper);
        double G = net_rate(u, nH, nd.XH, nd.rf);
        if (u - u_old - ratefact * G * dt > 0.0) u_upper = u;
        else u_lower = u;
        du = fabs(u_upper - u_lower);
        ++niter;
    }
    if (niter >= nd.max_iter) *capped = 1;
    *iters = niter - 1;
    return u;
}

// The network step of a particle in code units.
__device__ __forceinline__ double net_particle(double rho, double u, double dt, const NetDev &nd,
                                               int *iters, int *capped) {
    double un = net_cool_one(rho * nd.unit_rho_cgs, u * nd.unit_u_cgs, dt * nd.unit_time_s, nd,
                             iters, capped);
    return un / nd.unit_u_cgs;
}

__global__ void k_net(NetDev nd, const double *__restrict__ rho, const double *u_in, double dt,
                      int64_t n, double *u_out, int32_t *__restrict__ iters, int64_t *counters) {
    int64_t i = (int64_t)blockIdx.x * IT + threadIdx.x;
    if (i >= n) return;
    int it = 0, cap = 0;
    u_out[i] = net_particle(rho[i], u_in[i], dt, nd, &it, &cap);
    if (iters) iters[i] = it;
    if (cap && counters) atomicAdd((unsigned long long *)&counters[CNT_COOL_CAPPED], 1ull);
}

__global__ void k_cool(CoolDev cd, const double *__restrict__ rho, const double *__restrict__ u_in,
                       double dt, const int32_t *__restrict__ run, const RunParams *__restrict__ rp,
                       int64_t n, double *__restrict__ u_out, int32_t *__restrict__ iters,
                       int64_t *counters) {
    int64_t i = (int64_t)blockIdx.x * IT + threadIdx.x;
    if (i >= n) return;
    double d = dt;
    if (rp) {
        const RunParams &r = rp[run[i]];
        if (!r.use_cool) {
            u_out[i] = u_in[i];
            if (iters) iters[i] = 0;
            return;
        }
        d = r.dt;
    }
    int it = 0;
    u_out[i] = cool_one(rho[i], u_in[i], d, cd, &it);
    if (iters) iters[i] = it;
    if (it >= cd.max_iter && counters)
        atomicAdd((unsigned long long *)&counters[CNT_COOL_CAPPED], 1ull);
}

// Second half kick, u trapezoid (myCppSPHLibs_v2.h:740-747), ut_prev update,
// then the implicit cooling of the adiabatic u (operator split).
// First bin j of an increasing grid with g[j] <= x <= g[j+1] for g[0] <= x:
// the reference's linear scan (myCppSPHLibs_v3.h:831-849) as a binary search
// for the first k >= 1 with g[k] >= x; past the last node -> ng - 2.
__device__ __forceinline__ int64_t grid_bin(const double *__restrict__ g, int64_t ng, double x) {
    int64_t lo = 1, hi = ng;  // first k in [1, ng) with g[k] >= x, else ng
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (g[mid] >= x) hi = mid;
        else lo = mid + 1;
    }
    int64_t j = lo - 1;
    return j > ng - 2 ? ng - 2 : j;
}

// applyCooling of the 2-D table (myCppSPHLibs_v3.h:794-859), FP64, floored at
// the first u node.
__device__ double grid_apply(double rho, double u_code, double dt, const GridDev &g) {
#pragma clang fp contract(off)
    double coeff = (dt * g.unit_time_s) / g.dt_ref_s;
    double umin = g.u_nodes[0], umax = g.u_nodes[g.nu - 1];
    double rmin = g.rho_nodes[0], rmax = g.rho_nodes[g.nrho - 1];
    double uc = u_code * g.unit_u_cgs, rc = rho * g.unit_rho_cgs;
    if (uc < umin) uc = umin + 0.001 * umin;
    if (uc > umax) uc = umax - 0.001 * umax;
    if (rc < rmin) rc = rmin + 0.001 * rmin;
    if (rc > rmax) rc = rmax - 0.001 * rmax;
    int64_t ju = grid_bin(g.u_nodes, g.nu, uc), jr = grid_bin(g.rho_nodes, g.nrho, rc);
    double un = g.u_nodes[ju] - coeff * g.du[ju * g.nrho + jr];
    if (un < umin) un = umin;
    return un / g.unit_u_cgs;
}

__global__ void k_grid_apply(GridDev g, const double *__restrict__ rho,
                             const double *u_in, double dt, int64_t n, double *u_out) {
    int64_t i = (int64_t)blockIdx.x * IT + threadIdx.x;
    if (i < n) u_out[i] = grid_apply(rho[i], u_in[i], dt, g);
}

// du at every node from the implicit solve over dt_ref_s (cgs in and out).
__global__ void k_grid_build(GridDev g, CoolDev cd, double *__restrict__ du) {
#pragma clang fp contract(off)
    int64_t t = (int64_t)blockIdx.x * IT + threadIdx.x;
    if (t >= g.nu * g.nrho) return;
    int64_t j = t / g.nrho, i = t - j * g.nrho;
    double u_code = g.u_nodes[j] / cd.unit_u_cgs;
    double rho_code = g.rho_nodes[i] / cd.unit_rho_cgs;
    double dt_code = g.dt_ref_s / cd.unit_time_s;
    int it = 0;
    double un = cool_one(rho_code, u_code, dt_code, cd, &it);
    du[t] = (u_code - un) * cd.unit_u_cgs;
}

// hcoolingx of the 5-D CHIMES table (hfvCLibs_v7.h:1077-1240), FP64 on the
// float32 table.  Sorted axes are searched by bisection for the first node >=
// the value (the reference's linear scans find the same node); the T column
// is scanned in order, since its node energies need not increase.
__device__ __forceinline__ int chimes_first_ge(const float *__restrict__ g, int n, double x) {
    int lo = 0, hi = n;  // first k with x <= g[k], else n (also for NaN)
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (x <= (double)g[mid]) hi = mid;
        else lo = mid + 1;
    }
    return lo < n ? lo : -1;
}

__device__ __forceinline__ int chimes_nearest(const float *__restrict__ g, int n, double x) {
    int j = chimes_first_ge(g, n, x);
    if (j > 0) {
        double d1 = fabs((double)g[j] - x), d2 = fabs((double)g[j - 1] - x);
        if (d2 < d1) j -= 1;
    }
    return j < 0 ? n - 1 : j;
}

// New energy (code units) of one particle; *cell receives the flat index of
// the chosen (kpc, T, nH, Z, time) node in u_evol / mu and *mcell that of
// species 0 in metalz (species k at + k * n_time).
__device__ double chimes_one(double x, double y, double z, double rho, double u_code, double dt,
                             const ChimesDev &c, int64_t *cell, int64_t *mcell) {
#pragma clang fp contract(off)
    const int64_t NT = c.n_t, NH = c.n_nh, NZ = c.n_z, NM = c.n_m, NTIME = c.n_time;
    const int64_t s_t = NH * NZ * NTIME, s_k = NT * s_t;
    double dt_s = dt * c.unit_time_s;
    int jt = chimes_first_ge(c.time_s, c.n_time, dt_s);
    if (jt < 0) jt = c.n_time - 1;
    double a = x * x, b = y * y, cz = z * z;
    double r = sqrt((a + b) + cz);
    double r_kpc = (r * c.unit_length_cm) / c.kpc_cm;
    int64_t jk = chimes_nearest(c.kpc, c.n_kpc, r_kpc);
    double nh_i = ((rho * c.unit_rho_cgs) * c.xh) / c.m_h;
    int64_t jn = chimes_nearest(c.nh, c.n_nh, nh_i);
    int64_t jz = c.z_index;
    int64_t col = jk * s_k + jn * NZ * NTIME + jz * NTIME;
    double u_i = u_code * c.unit_u_cgs;
    int64_t ju = NT - 1;
    for (int64_t t = 0; t < NT; ++t)
        if (u_i <= (double)c.u_evol[col + t * s_t]) { ju = t; break; }
    double uev;
    if (ju > 0) {
        double U0 = c.u_evol[col + (ju - 1) * s_t], U1 = c.u_evol[col + ju * s_t];
        double e0 = c.u_evol[col + (ju - 1) * s_t + jt], e1 = c.u_evol[col + ju * s_t + jt];
        double fhi = (u_i - U0) / (U1 - U0);
        double flo = 1.0 - fhi;
        uev = flo * e0 + fhi * e1;
    } else {
        uev = c.u_evol[col + jt];
    }
    if (uev > c.u_max_cgs) uev = c.u_max_cgs;
    *cell = col + ju * s_t + jt;
    *mcell = ((jk * NT + ju) * NH + jn) * NZ * NM * NTIME + jz * NM * NTIME + jt;
    return uev / c.unit_u_cgs;
}

__global__ void k_chimes_apply(ChimesDev c, const double *__restrict__ pos,
                               const double *__restrict__ rho, const double *u_in, double dt,
                               int64_t n, double *u_out, float *__restrict__ ion_out,
                               double *__restrict__ mu_out) {
    int64_t i = (int64_t)blockIdx.x * IT + threadIdx.x;
    if (i >= n) return;
    int64_t cell, mcell;
    double un = chimes_one(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], rho[i], u_in[i], dt, c,
                           &cell, &mcell);
    u_out[i] = un;
    if (mu_out) mu_out[i] = c.mu[cell];
    if (ion_out)
        for (int k = 0; k < c.n_m; ++k)
            ion_out[i * c.n_m + k] = c.metalz[mcell + (int64_t)k * c.n_time];
}

// Cooling of particle i's updated energy u over dt by the run's method (0
// none, 1 implicit solve, 2 2-D table, 3 CHIMES table).
// *iters receives the bisection iterations of an implicit solve (0 otherwise).
__device__ __forceinline__ double cool_particle(const Soa &s, int64_t i, double u, double dt,
                                                int use_cool, int cool_on, const CoolDev &cd,
                                                const GridDev &gd, const ChimesDev &ch,
                                                const NetDev &nd, int64_t *counters, int *iters) {
    *iters = 0;
    if (cool_on && use_cool == 1) {
        int it = 0;
        u = cool_one(s.rho[i], u, dt, cd, &it);
        *iters = it;
        if (it >= cd.max_iter) atomicAdd((unsigned long long *)&counters[CNT_COOL_CAPPED], 1ull);
    } else if (use_cool == 4 && nd.on) {
        int it = 0, cap = 0;
        u = net_particle(s.rho[i], u, dt, nd, &it, &cap);
        *iters = it;
        if (cap) atomicAdd((unsigned long long *)&counters[CNT_COOL_CAPPED], 1ull);
    } else if (use_cool == 2 && gd.du) {
        u = grid_apply(s.rho[i], u, dt, gd);
    } else if (use_cool == 3 && ch.u_evol) {
        int64_t cell, mcell;
        u = chimes_one(s.x[i], s.y[i], s.z[i], s.rho[i], u, dt, ch, &cell, &mcell);
        if (ch.ion) {
            float *o = ch.ion + (int64_t)s.id[i] * ch.n_m;
            for (int k = 0; k < ch.n_m; ++k) o[k] = ch.metalz[mcell + (int64_t)k * ch.n_time];
        }
    }
    return u;
}

// Adds the bisection iterations of a wavefront's cooling solves and the
// number of its lanes that bisected to the step counters, one atomic each per
// wavefront; every lane of the wavefront must call it.
__device__ __forceinline__ void count_cool_iters(int it, int64_t *counters) {
    unsigned long long sum = (unsigned long long)it;
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
    unsigned long long nb = (unsigned long long)__popcll(__ballot(it > 0));
    if ((threadIdx.x & 63) == 0 && nb) {
        atomicAdd((unsigned long long *)&counters[CNT_COOL_ITERS], sum);
        atomicAdd((unsigned long long *)&counters[CNT_COOL_BISECT], nb);
    }
}

// A particle whose state stopped being finite is counted (the step goes on;
// the caller reads the count), as the reference has no such check at all.
__device__ __forceinline__ void count_nonfinite(const Soa &s, int64_t i, double u,
                                                int64_t *counters) {
    if (!(isfinite(u) && isfinite(s.vx[i]) && isfinite(s.vy[i]) && isfinite(s.vz[i]) &&
          isfinite(s.x[i]) && isfinite(s.y[i]) && isfinite(s.z[i])))
        atomicAdd((unsigned long long *)&counters[CNT_NONFINITE], 1ull);
}

__global__ void k_finish(Soa s, const RunParams *__restrict__ rp, int64_t n, CoolDev cd,
                         int cool_on, GridDev gd, ChimesDev ch, NetDev nd, int64_t *counters) {
#pragma clang fp contract(off)
    int64_t i = (int64_t)blockIdx.x * IT + threadIdx.x;
    int it = 0;
    if (i < n) {
        const RunParams &r = rp[s.run[i]];
        double dt = r.dt;
        double half = 0.5 * dt;
        s.vx[i] = s.vx[i] + s.ax[i] * half;
        s.vy[i] = s.vy[i] + s.ay[i] * half;
        s.vz[i] = s.vz[i] + s.az[i] * half;
        double du = s.du[i];
        double u = s.u[i] + half * (du + s.ut_prev[i]);
        s.ut_prev[i] = du;
        u = cool_particle(s, i, u, dt, r.use_cool, cool_on, cd, gd, ch, nd, counters, &it);
        s.u[i] = u;
        count_nonfinite(s, i, u, counters);
    }
    count_cool_iters(it, counters);
}

// ---------------------------------------------------------------------------
// Block time steps: particle i steps with dt_max / 2^bin[i]; the substep is
// dt_max / 2^B.  A particle is active at the end of a substep (tick t1, in
// substeps) when t1 is a multiple of its step.  Kick-drift-kick per particle:
// the opening half kick when its step starts, every particle drifts every
// substep, forces at the end of its step, the closing half kick, a new bin.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool blk_active(int b, int B, int64_t t1) {
    return (t1 & ((1ll << (B - b)) - 1)) == 0;
}

// The per-particle form of the step criterion of the global dt
// (myCppSPHLibs_v2.h:491-522 with hfv_sph_gpu_v2_evrard.cu:555-561: the same
// four limits, each from this particle's own h, c, |a|, |a_g|, div v).
__device__ __forceinline__ double blk_dt_crit(const Soa &s, int64_t i, double c_cfl) {
#pragma clang fp contract(off)
    double h = s.h[i];
    double ag = sqrt((s.gx[i] * s.gx[i] + s.gy[i] * s.gy[i]) + s.gz[i] * s.gz[i]);
    double at = sqrt((s.ax[i] * s.ax[i] + s.ay[i] * s.ay[i]) + s.az[i] * s.az[i]);
    double dd = h / fabs((1.0 / 3.0) * h * s.divv[i]);
    double dt_f = sqrt(h / ag), dt_kin = sqrt(h / at);
    double dt_cfl = c_cfl * h / s.c[i], dt_dens = c_cfl * dd;
    double m = dt_f;
    m = dt_kin < m ? dt_kin : m;
    m = dt_cfl < m ? dt_cfl : m;
    m = dt_dens < m ? dt_dens : m;
    return 0.25 * m;
}

// The longest step dt_max / 2^b (b in 0..B) not above dt_i; B (and a floor
// count) when even the shortest is too long or dt_i is not a number.
__device__ __forceinline__ int blk_target(double dt_i, double dt_max, int B, bool *floor) {
    for (int b = 0; b <= B; ++b)
        if (ldexp(dt_max, -b) <= dt_i) return b;
    *floor = true;
    return B;
}

// Tick 0: every particle takes its first bin and its opening half kick.
__global__ void k_blk_open(Soa s, int64_t n, double dt_max, int B, double c_cfl,
                           int64_t *counters) {
#pragma clang fp contract(off)
    int64_t i = (int64_t)blockIdx.x * IT + threadIdx.x;
    if (i >= n) return;
    bool fl = false;
    int b = blk_target(blk_dt_crit(s, i, c_cfl), dt_max, B, &fl);
    s.bin[i] = b;
    double half = 0.5 * ldexp(dt_max, -b);
    s.vx[i] = s.vx[i] + s.ax[i] * half;
    s.vy[i] = s.vy[i] + s.ay[i] * half;
    s.vz[i] = s.vz[i] + s.az[i] * half;
    if (fl) atomicAdd((unsigned long long *)&counters[CNT_BIN_FLOOR], 1ull);
}

__global__ void k_blk_drift(Soa s, int64_t n, double d) {
#pragma clang fp contract(off)
    int64_t i = (int64_t)blockIdx.x * IT + threadIdx.x;
    if (i >= n) return;
    s.x[i] = s.x[i] + s.vx[i] * d;
    s.y[i] = s.y[i] + s.vy[i] * d;
    s.z[i] = s.z[i] + s.vz[i] * d;
}

// End of substep t1: the active particles take the closing half kick, the u
// trapezoid and cooling over their step, a new bin (a shorter step at once, a
// longer one only one level at a time and only where t1 is a multiple of the
// longer step), then the opening half kick of their next step.  This is one
// active particle i of k_blk_close.
__device__ __forceinline__ void blk_close_one(Soa &s, const RunParams *__restrict__ rp, int64_t i,
                                              double dt_max, int B, int64_t t1, double c_cfl,
                                              const CoolDev &cd, int cool_on, const GridDev &gd,
                                              const ChimesDev &ch, const NetDev &nd,
                                              int64_t *counters, int *it) {
#pragma clang fp contract(off)
    int b = s.bin[i];
    double dto = ldexp(dt_max, -b);
    double half = 0.5 * dto;
    s.vx[i] = s.vx[i] + s.ax[i] * half;
    s.vy[i] = s.vy[i] + s.ay[i] * half;
    s.vz[i] = s.vz[i] + s.az[i] * half;
    double du = s.du[i];
    double u = s.u[i] + half * (du + s.ut_prev[i]);
    s.ut_prev[i] = du;
    u = cool_particle(s, i, u, dto, rp[s.run[i]].use_cool, cool_on, cd, gd, ch, nd, counters, it);
    s.u[i] = u;
    bool fl = false;
    int bt = blk_target(blk_dt_crit(s, i, c_cfl), dt_max, B, &fl);
    int bn = b;
    if (bt >= b) bn = bt;
    else if (blk_active(b - 1, B, t1)) bn = b - 1;
    s.bin[i] = bn;
    double h2 = 0.5 * ldexp(dt_max, -bn);
    s.vx[i] = s.vx[i] + s.ax[i] * h2;
    s.vy[i] = s.vy[i] + s.ay[i] * h2;
    s.vz[i] = s.vz[i] + s.az[i] * h2;
    if (fl) atomicAdd((unsigned long long *)&counters[CNT_BIN_FLOOR], 1ull);
    count_nonfinite(s, i, u, counters);
}

__global__ void k_blk_close(Soa s, const RunParams *__restrict__ rp, int64_t n, double dt_max,
                            int B, int64_t t1, double c_cfl, CoolDev cd, int cool_on, GridDev gd,
                            ChimesDev ch, NetDev nd, int64_t *counters) {
#pragma clang fp contract(off)
    int64_t i = (int64_t)blockIdx.x * IT + threadIdx.x;
    int it = 0;
    if (i < n && blk_active(s.bin[i], B, t1))
        blk_close_one(s, rp, i, dt_max, B, t1, c_cfl, cd, cool_on, gd, ch, nd, counters, &it);
    count_cool_iters(it, counters);
}

__global__ void k_copy_d(const double *__restrict__ a, double *__restrict__ b, int64_t p0,
                         int64_t p1) {
    int64_t i = p0 + (int64_t)blockIdx.x * IT + threadIdx.x;
    if (i < p1) b[i] = a[i];
}

// Per-block partial sums of the 11 diagnostic quantities; fixed tree order.
constexpr int ND = 11;
__global__ __launch_bounds__(IT) void k_diag_partial(const double *x, const double *y,
                                                     const double *z, const double *vx,
                                                     const double *vy, const double *vz,
                                                     const double *m, const double *u,
                                                     const double *pot, int64_t n,
                                                     double *__restrict__ part) {
    __shared__ double sh[ND][IT];
    int tid = threadIdx.x;
    double v[ND];
    for (int a = 0; a < ND; ++a) v[a] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * IT + tid; i < n; i += (int64_t)gridDim.x * IT) {
        double mi = m[i];
        double a = vx[i], b = vy[i], c = vz[i];
        v[0] += 0.5 * mi * (a * a + b * b + c * c);
        v[1] += u ? mi * u[i] : 0.0;
        v[2] += pot ? 0.5 * mi * pot[i] : 0.0;
        v[3] += mi * a;
        v[4] += mi * b;
        v[5] += mi * c;
        if (x) {
            double px = x[i], py = y[i], pz = z[i];
            v[6] += mi * (py * c - pz * b);
            v[7] += mi * (pz * a - px * c);
            v[8] += mi * (px * b - py * a);
        }
        v[9] += mi;
    }
    for (int a = 0; a < ND; ++a) sh[a][tid] = v[a];
    __syncthreads();
    for (int off = IT / 2; off > 0; off >>= 1) {
        if (tid < off)
            for (int a = 0; a < ND; ++a) sh[a][tid] += sh[a][tid + off];
        __syncthreads();
    }
    if (tid < ND) part[(int64_t)blockIdx.x * ND + tid] = sh[tid][0];
}

}  // namespace

int sphx_cool_launch(const CoolDev &cd, const double *rho, const double *u_in, double dt,
                     const double *, const int32_t *run, const RunParams *rp, int64_t n,
                     double *u_out, int32_t *iters, int64_t *counters, hipStream_t st) {
    if (n <= 0) return SPHX_OK;
    k_cool<<<g1(n), IT, 0, st>>>(cd, rho, u_in, dt, run, rp, n, u_out, iters, counters);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int sphx_grid_apply_launch(const GridDev &g, const double *rho, const double *u_in, double dt,
                           int64_t n, double *u_out, hipStream_t st) {
    if (n <= 0) return SPHX_OK;
    k_grid_apply<<<g1(n), IT, 0, st>>>(g, rho, u_in, dt, n, u_out);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int sphx_chimes_apply_launch(const ChimesDev &c, const double *pos, const double *rho,
                             const double *u_in, double dt, int64_t n, double *u_out,
                             float *ion_out, double *mu_out, hipStream_t st) {
    if (n <= 0) return SPHX_OK;
    k_chimes_apply<<<g1(n), IT, 0, st>>>(c, pos, rho, u_in, dt, n, u_out, ion_out, mu_out);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int sphx_grid_build_launch(const GridDev &g, const CoolDev &cd, double *du, hipStream_t st) {
    int64_t nn = g.nu * g.nrho;
    if (nn <= 0) return SPHX_OK;
    k_grid_build<<<g1(nn), IT, 0, st>>>(g, cd, du);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int sphx_diag_reduce(const double *x, const double *y, const double *z, const double *vx,
                     const double *vy, const double *vz, const double *m, const double *u,
                     const double *pot, int64_t n, double *scratch, sphx_diag *out,
                     hipStream_t st) {
    const int blocks = 256;
    k_diag_partial<<<blocks, IT, 0, st>>>(x, y, z, vx, vy, vz, m, u, pot, n, scratch);
    SPHX_HIP(hipGetLastError());
    std::vector<double> h((size_t)blocks * ND);
    SPHX_HIP(hipMemcpyAsync(h.data(), scratch, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    SPHX_HIP(hipStreamSynchronize(st));
    double v[ND] = {0};
    // pairwise combination of the block partials in a fixed order
    std::vector<double> cur(h);
    int len = blocks;
    while (len > 1) {
        int half = (len + 1) / 2;
        for (int b = 0; b < len / 2; ++b)
            for (int a = 0; a < ND; ++a) cur[(size_t)b * ND + a] += cur[(size_t)(b + half) * ND + a];
        len = half;
    }
    for (int a = 0; a < ND; ++a) v[a] = cur[a];
    out->ke = v[0]; out->ut = v[1]; out->pe = v[2];
    out->px = v[3]; out->py = v[4]; out->pz = v[5];
    out->lx = v[6]; out->ly = v[7]; out->lz = v[8];
    out->mass = v[9];
    return SPHX_OK;
}

int Engine::kick_drift(hipStream_t st) {
    if (n == 0) return SPHX_OK;
    k_kick_drift<<<g1(n), IT, 0, st>>>(s, d_run, n);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

static CoolDev make_cool(const Engine &E) {
    CoolDev cd{};
    if (E.has_cool) {
        const sphx_cooling &c = E.cool;
        cd = CoolDev{c.unit_u_cgs, c.unit_rho_cgs, c.unit_time_s, c.XH, c.mu, c.gamma, c.heat,
                     c.tol, c.max_iter, c.max_bracket, c.ntab, E.cool_T, E.cool_L};
    }
    return cd;
}

CoolDev sphx_engine_cool(const Engine &E) { return make_cool(E); }

int Engine::finish_step(hipStream_t st) {
    if (n == 0) return SPHX_OK;
    k_finish<<<g1(n), IT, 0, st>>>(s, d_run, n, make_cool(*this), has_cool ? 1 : 0, grid,
                                   chimes, net, counters);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int Engine::block_open(hipStream_t st) {
    if (n == 0) return SPHX_OK;
    k_blk_open<<<g1(n), IT, 0, st>>>(s, n, blk_dtmax, blk_B, blk_c_cfl, counters);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int Engine::block_drift(double d, hipStream_t st) {
    if (n == 0) return SPHX_OK;
    k_blk_drift<<<g1(n), IT, 0, st>>>(s, n, d);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int Engine::block_close(int64_t t1, hipStream_t st) {
    if (n == 0) return SPHX_OK;
    k_blk_close<<<g1(n), IT, 0, st>>>(s, d_run, n, blk_dtmax, blk_B, t1, blk_c_cfl,
                                      make_cool(*this), has_cool ? 1 : 0, grid, chimes, net,
                                      counters);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int Engine::bootstrap_finish(hipStream_t st) {
    if (n == 0) return SPHX_OK;
    int64_t p0 = rrun >= 0 ? rp0 : 0, p1 = rrun >= 0 ? rp1 : n;
    if (p1 <= p0) return SPHX_OK;
    k_copy_d<<<g1(p1 - p0), IT, 0, st>>>(s.du, s.ut_prev, p0, p1);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int sphx_zero_counters(int64_t *c, hipStream_t st) {
    k_zero_i64<<<1, 64, 0, st>>>(c, CNT_SLOTS);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int sphx_combine(Engine &E, hipStream_t st) {
    if (E.n == 0) return SPHX_OK;
    int64_t p0 = E.rrun >= 0 ? E.rp0 : 0, p1 = E.rrun >= 0 ? E.rp1 : E.n;
    if (p1 <= p0) return SPHX_OK;
    k_combine<<<g1(p1 - p0), IT, 0, st>>>(E.s, p0, p1);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

// ---------------------------------------------------------------------------
// Global time-step criterion (myCppSPHLibs_v2.h:491-522 dt_array and the
// host reduction of hfv_sph_gpu_v2_evrard.cu:517-561), reduced on the device:
// per 64-particle bucket the five extrema (max c, min h, max |a_grav|,
// max |a_total|, min h/|h divV/3|), then per run the four limits
//   dt_f = sqrt(hmin/max|a_g|), dt_kin = sqrt(hmin/max|a|),
//   dt_cfl = C_CFL hmin / max c, dt_dens = C_CFL min h/|h divV/3|,
// dt = 0.25 min(...) clamped to [dt_min, dt_max].  The reference's O(N^2)
// v_sig loop (max_j max(c_i, c_j)) is the global max of c.  min/max are
// order-independent, so the result is bit-identical to a serial FP64 pass.
// ---------------------------------------------------------------------------
namespace {

struct DtExt {
    double v[5];  // vmax, hmin, agmax, amax, ddmin
};

__device__ __forceinline__ void dt_ident(DtExt &e) {
    e.v[0] = 0.0; e.v[1] = INFINITY; e.v[2] = 0.0; e.v[3] = 0.0; e.v[4] = INFINITY;
}

__device__ __forceinline__ void dt_merge(DtExt &e, const DtExt &o) {
    e.v[0] = fmax(e.v[0], o.v[0]);
    e.v[1] = fmin(e.v[1], o.v[1]);
    e.v[2] = fmax(e.v[2], o.v[2]);
    e.v[3] = fmax(e.v[3], o.v[3]);
    e.v[4] = fmin(e.v[4], o.v[4]);
}

__device__ __forceinline__ void dt_terms(double h, double c, double ax, double ay, double az,
                                         double gx, double gy, double gz, double dv, DtExt &e) {
#pragma clang fp contract(off)
    e.v[0] = c;
    e.v[1] = h;
    e.v[2] = sqrt((gx * gx + gy * gy) + gz * gz);
    e.v[3] = sqrt((ax * ax + ay * ay) + az * az);
    e.v[4] = h / fabs((1.0 / 3.0) * h * dv);
}

__device__ __forceinline__ void dt_wave(DtExt &e) {
    for (int o = 32; o > 0; o >>= 1) {
        DtExt t;
        for (int k = 0; k < 5; ++k) t.v[k] = __shfl_xor(e.v[k], o, 64);
        dt_merge(e, t);
    }
}

// out: dt, dt_f, dt_kin, dt_cfl, dt_dens
__device__ __forceinline__ void dt_limits(const DtExt &e, double c_cfl, double dt_min,
                                          double dt_max, double *out) {
#pragma clang fp contract(off)
    double dt_f = sqrt(e.v[1] / e.v[2]), dt_kin = sqrt(e.v[1] / e.v[3]);
    double dt_cfl = c_cfl * e.v[1] / e.v[0], dt_dens = c_cfl * e.v[4];
    double m = dt_f;
    m = dt_kin < m ? dt_kin : m;
    m = dt_cfl < m ? dt_cfl : m;
    m = dt_dens < m ? dt_dens : m;
    double dt = 0.25 * m;
    if (dt > dt_max) dt = dt_max;
    if (dt < dt_min) dt = dt_min;
    out[0] = dt; out[1] = dt_f; out[2] = dt_kin; out[3] = dt_cfl; out[4] = dt_dens;
}

// One wavefront per bucket (b_beg == nullptr: bucket b is [64 b, 64 b + 64)).
__global__ __launch_bounds__(64) void k_dt_part(const double *__restrict__ h,
                                                const double *__restrict__ c,
                                                const double *__restrict__ ax,
                                                const double *__restrict__ ay,
                                                const double *__restrict__ az, int64_t as,
                                                const double *__restrict__ gx,
                                                const double *__restrict__ gy,
                                                const double *__restrict__ gz, int64_t gs,
                                                const double *__restrict__ dv,
                                                const int32_t *__restrict__ b_beg,
                                                const int32_t *__restrict__ b_cnt, int64_t n,
                                                double *__restrict__ part) {
    const int64_t b = blockIdx.x;
    int64_t beg = b_beg ? (int64_t)b_beg[b] : 64 * b;
    int64_t cnt = b_beg ? (int64_t)b_cnt[b] : (n - beg < 64 ? n - beg : 64);
    DtExt e;
    dt_ident(e);
    if ((int64_t)threadIdx.x < cnt) {
        int64_t i = beg + threadIdx.x;
        dt_terms(h[i], c[i], ax[as * i], ay[as * i], az[as * i], gx[gs * i], gy[gs * i], gz[gs * i],
                 dv[i], e);
    }
    dt_wave(e);
    if (threadIdx.x == 0)
        for (int k = 0; k < 5; ++k) part[5 * b + k] = e.v[k];
}

__device__ void dt_block(const double *__restrict__ part, int64_t b0, int64_t b1, DtExt &e) {
    __shared__ double red[4][5];
    dt_ident(e);
    for (int64_t b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
        DtExt o;
        for (int k = 0; k < 5; ++k) o.v[k] = part[5 * b + k];
        dt_merge(e, o);
    }
    dt_wave(e);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 5; ++k) red[w][k] = e.v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
            DtExt o;
            for (int k = 0; k < 5; ++k) o.v[k] = red[q][k];
            dt_merge(e, o);
        }
}

// One workgroup per run: t += dt; adaptive runs take the new dt.
__global__ __launch_bounds__(256) void k_dt_runs(const double *__restrict__ part, Tables t,
                                                 RunParams *__restrict__ rp) {
    const int r = blockIdx.x;
    int64_t b0 = 0, b1 = 0;
    if (t.r_send[r] > t.r_sbeg[r]) {
        b0 = t.s_bbeg[t.r_sbeg[r]];
        b1 = t.s_bend[t.r_send[r] - 1];
    }
    DtExt e;
    dt_block(part, b0, b1, e);
    if (threadIdx.x == 0) {
        RunParams q = rp[r];
        q.t = q.t + q.dt;
        if (b1 > b0) {
            double o[5];
            dt_limits(e, q.c_cfl, q.dt_min, q.dt_max, o);
            for (int k = 0; k < 4; ++k) q.dt_parts[k] = o[k + 1];
            if (q.adaptive) q.dt = o[0];
        }
        rp[r] = q;
    }
}

__global__ __launch_bounds__(256) void k_dt_one(const double *__restrict__ part, int64_t nb,
                                                double c_cfl, double dt_min, double dt_max,
                                                double *__restrict__ out) {
    DtExt e;
    dt_block(part, 0, nb, e);
    if (threadIdx.x == 0) dt_limits(e, c_cfl, dt_min, dt_max, out);
}

}  // namespace

int Engine::global_dt(hipStream_t st) {
    if (n == 0 || R == 0) return SPHX_OK;
    k_dt_part<<<(unsigned)t.nb, 64, 0, st>>>(s.h, s.c, s.ax, s.ay, s.az, 1, s.gx, s.gy, s.gz, 1,
                                             s.divv, t.b_beg, t.b_cnt, n, dtpart);
    k_dt_runs<<<(unsigned)R, 256, 0, st>>>(dtpart, t, d_run);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int sphx_global_dt_launch(const double *h, const double *c, const double *acc, int64_t as,
                          const double *acc_g, int64_t gs, const double *divv, int64_t n,
                          double c_cfl, double dt_min, double dt_max, double *scratch,
                          int64_t scratch_len, double *out, hipStream_t st) {
    int64_t nb = (n + 63) / 64;
    if (n <= 0 || 5 * nb > scratch_len) {
        sphx_set_error("global dt: bad size or scratch too small");
        return SPHX_ERR_ARG;
    }
    k_dt_part<<<(unsigned)nb, 64, 0, st>>>(h, c, acc, acc + 1, acc + 2, as, acc_g, acc_g + 1,
                                           acc_g + 2, gs, divv, nullptr, nullptr, n, scratch);
    k_dt_one<<<1, 256, 0, st>>>(scratch, nb, c_cfl, dt_min, dt_max, out);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

// ---------------------------------------------------------------------------
// Host side of the network: the background rates (RadiationField,
// photolibs3.py:230-329): for each ionisation edge (H0 13.60 eV, He0 24.59 eV,
// He+ 54.42 eV) the trapezoid sums over the frequency grid of
// 4 pi J sigma / (h nu) (photoionisation) and 4 pi J sigma (nu - nu_edge) / nu
// (heating), with the cross sections of Verner et al. (1996).
// ---------------------------------------------------------------------------
namespace {

constexpr double NET_H_EV = 4.135667696e-15;  // Planck constant in eV s
constexpr double NET_CLIGHT = 29979245800.0;  // cm/s

double net_sigma(double nu, const double *c) {
    double E = NET_H_EV * nu;
    double x = E / c[0] - c[5];
    double y = sqrt(x * x + c[6] * c[6]);
    double Fy = (((x - 1.0) * (x - 1.0) + c[4] * c[4]) * pow(y, 0.5 * c[3] - 5.5)) *
                pow(1.0 + sqrt(y / c[2]), -c[3]);
    return c[1] * Fy * 1e-18;
}

void net_edge(const sphx_network &p, double E, const double *c, double *g, double *H) {
    const double hplanck = 6.626196e-27;
    const double nu0 = E / NET_H_EV;
    const double start = nu0 / 1e16;
    int64_t n;
    double step;
    if (p.n_grid > 0) {
        n = p.n_grid;
        step = (6.0 - start) / (double)n;
    } else {  // the values of np.arange(start, 6.0, 0.02)
        n = (int64_t)ceil((6.0 - start) / 0.02);
        step = (start + 0.02) - start;
    }
    // the edge frequency through wavelength, as the reference's J_neu_func
    const double lamb = NET_CLIGHT / nu0 * 1e8;
    const double nu_edge = NET_CLIGHT / (lamb * 1e-8);
    std::vector<double> fg(n), fh(n), nu(n);
    for (int64_t i = 0; i < n; ++i) {
        nu[i] = (start + (double)i * step) * 1e16;
        double sig = net_sigma(nu[i], c);
        double J = p.J21 * 1e-21 / (nu[i] / nu_edge);
        fg[i] = 4.0 * M_PI * J * sig / hplanck / nu[i];
        fh[i] = 4.0 * M_PI * J * sig * (nu[i] - nu0) / nu[i];
    }
    double d = n > 1 ? nu[1] - nu[0] : 0.0;
    double sg = 0.0, sh = 0.0;
    for (int64_t i = 0; i + 1 < n; ++i) {
        sg += d * (fg[i] + fg[i + 1]) / 2.0;
        sh += d * (fh[i] + fh[i + 1]) / 2.0;
    }
    *g = sg;
    *H = sh;
}

}  // namespace

extern "C" int sphx_network_rates(const sphx_network *net, double *out) {
    if (!net || !out) { sphx_set_error("null argument"); return SPHX_ERR_ARG; }
    if (!(net->J21 >= 0.0) || net->n_grid < 0 || net->n_grid > 1000000) {
        sphx_set_error("sphx_network: J21 >= 0 and 0 <= n_grid <= 1e6");
        return SPHX_ERR_ARG;
    }
    // Verner et al. (1996) parameters E0, sigma0, ya, P, yw, y0, y1
    static const double cH0[7] = {4.298e-1, 5.475e4, 3.288e1, 2.963, 0.0, 0.0, 0.0};
    static const double cHe0[7] = {1.361e1, 9.492e2, 1.469e0, 3.188, 2.039, 4.434e-1, 2.136};
    static const double cHep[7] = {1.720e0, 1.369e4, 3.288e1, 2.963, 0.0, 0.0, 0.0};
    net_edge(*net, 13.60, cH0, &out[0], &out[3]);
    net_edge(*net, 24.59, cHe0, &out[1], &out[4]);
    net_edge(*net, 54.42, cHep, &out[2], &out[5]);
    return SPHX_OK;
}

int sphx_network_dev(const sphx_network *net, NetDev *out) {
    if (!net || !out) { sphx_set_error("null argument"); return SPHX_ERR_ARG; }
    if (!(net->XH > 0.0 && net->XH <= 1.0) || !(net->unit_u_cgs > 0) ||
        !(net->unit_rho_cgs > 0) || !(net->unit_time_s > 0) || !(net->tol > 0) ||
        net->max_iter <= 0 || net->max_bracket < 0) {
        sphx_set_error("invalid sphx_network (XH in (0, 1], positive units and tol, max_iter > 0)");
        return SPHX_ERR_ARG;
    }
    NetDev d;
    SPHX_CHECK(sphx_network_rates(net, d.rf));
    d.XH = net->XH;
    d.tol = net->tol;
    d.unit_u_cgs = net->unit_u_cgs;
    d.unit_rho_cgs = net->unit_rho_cgs;
    d.unit_time_s = net->unit_time_s;
    d.max_iter = net->max_iter;
    d.max_bracket = net->max_bracket;
    d.on = 1;
    *out = d;
    return SPHX_OK;
}

int sphx_net_launch(const NetDev &nd, const double *rho, const double *u_in, double dt, int64_t n,
                    double *u_out, int32_t *iters, int64_t *counters, hipStream_t st) {
    if (n <= 0) return SPHX_OK;
    k_net<<<g1(n), IT, 0, st>>>(nd, rho, u_in, dt, n, u_out, iters, counters);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free global time step: the kernels of sphx_global_dt with their
// per-bucket extrema in a fixed device array instead of a context's buffer.
// Each device holds its own copy of the array until the process ends; its
// address is fetched once per device with the device current, and only that
// address is handed to kernels.  Nothing in the array needs clearing: every
// slot a call reads was written by the same call.  Calls that overlap on one
// device share the array and race.
// ---------------------------------------------------------------------------
namespace {

constexpr int64_t DT_NOALLOC_MAX_N = int64_t(1) << 23;                  // 8,388,608 particles
constexpr int64_t DT_NOALLOC_SLOTS = 5 * ((DT_NOALLOC_MAX_N + 63) / 64);  // 5 MiB of doubles
constexpr int DT_NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[DT_NOALLOC_SLOTS];

std::mutex g_scratch_mu;
double *g_scratch_ptr[DT_NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on `device` (current), looked up on the first
// call per device.
int scratch_on(int device, double **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            sphx_set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return SPHX_ERR_HIP;
        }
        g_scratch_ptr[device] = (double *)p;
    }
    *out = g_scratch_ptr[device];
    return SPHX_OK;
}

}  // namespace

extern "C" int sphx_global_dt_noalloc(int device, int64_t n, const double *h, const double *c,
                                      const double *acc, const double *acc_g, const double *divV,
                                      double c_cfl, double dt_min, double dt_max, double *out,
                                      void *stream) {
    if (n <= 0 || !h || !c || !acc || !acc_g || !divV || !out) {
        sphx_set_error("bad arguments");
        return SPHX_ERR_ARG;
    }
    if (n > DT_NOALLOC_MAX_N) {
        sphx_set_error("sphx_global_dt_noalloc: n = " + std::to_string(n) +
                       " exceeds its fixed scratch (at most " +
                       std::to_string(DT_NOALLOC_MAX_N) + " particles); use sphx_global_dt");
        return SPHX_ERR_ARG;
    }
    if (!(c_cfl > 0.0 && dt_min > 0.0 && dt_max >= dt_min)) {
        sphx_set_error("global dt needs c_cfl > 0 and 0 < dt_min <= dt_max");
        return SPHX_ERR_ARG;
    }
    int ndev = 0;
    SPHX_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev || device >= DT_NOALLOC_MAX_DEVICES) {
        sphx_set_error("sphx_global_dt_noalloc: no such device " + std::to_string(device));
        return SPHX_ERR_ARG;
    }
    int prev = 0;
    SPHX_HIP(hipGetDevice(&prev));
    if (prev != device) SPHX_HIP(hipSetDevice(device));
    double *buf = nullptr;
    int rc = scratch_on(device, &buf);
    if (rc == SPHX_OK)
        rc = sphx_global_dt_launch(h, c, acc, 3, acc_g, 3, divV, n, c_cfl, dt_min, dt_max, buf,
                                   DT_NOALLOC_SLOTS, out, (hipStream_t)stream);
    if (prev != device) hipSetDevice(prev);
    return rc;
}
