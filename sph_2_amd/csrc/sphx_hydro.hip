// Bucketed SPH pair passes for the mixed-precision path: smoothing-length
// iteration, density + EOS, velocity divergence/curl, hydro acceleration +
// du/dt.  One wavefront per 64-particle bucket (one particle per lane); the
// candidate buckets of the same run are found by culling super-bucket and
// bucket AABBs in FP64, staged through LDS, and tested pair by pair with the
// oracle's FP64 distance.  Kernel values and pair sums run in FP32 with
// Kahan-compensated accumulation.
//
// Reference kernels replaced (11_Dec_2022/GPU_sph/Important_Test_Simulations/
// Evrard_collapse/myCppSPHLibs_v2.h): smoothing_h 7-64, getDensity 69-114,
// div_curlVel 247-322, acc_sph 327-417, get_dU 634-719,
// getPressure_Adiabatic/getCsound_Adiabatic 766-789.
#include "sphx_internal.h"

#include <algorithm>

namespace {

// Walk the candidate buckets, in increasing order, of a group of particles of
// run `r` with bounding box ib (xl, xh, yl, yh, zl, zh).  `reach_s(S)` and
// `reach_b(J)` give the reach against super S and bucket J; `body(J)` is
// called once per surviving bucket with the whole wavefront converged.
template <class ReachS, class ReachB, class Body>
__device__ __forceinline__ void for_candidates(const Tables &t, int r, const double *ib, int lane,
                                               ReachS reach_s, ReachB reach_b, Body body) {
    double axl = ib[0], axh = ib[1], ayl = ib[2], ayh = ib[3], azl = ib[4], azh = ib[5];
    int s0 = t.r_sbeg[r], s1 = t.r_send[r];
    for (int sb = s0; sb < s1; sb += 64) {
        int S = sb + lane;
        bool sp = false;
        if (S < s1) {
            double d2 = sphx_box_d2(axl, axh, ayl, ayh, azl, azh, t.sxl[S], t.sxh[S], t.syl[S],
                                    t.syh[S], t.szl[S], t.szh[S]);
            sp = sphx_within(d2, reach_s(S));
        }
        uint64_t smask = __ballot(sp);
        while (smask) {
            int S2 = sb + __builtin_ctzll(smask);
            smask &= smask - 1;
            int j0 = t.s_bbeg[S2], j1 = t.s_bend[S2];
            int J = j0 + lane;
            bool bp = false;
            if (J < j1) {
                double d2 = sphx_box_d2(axl, axh, ayl, ayh, azl, azh, t.bxl[J], t.bxh[J],
                                        t.byl[J], t.byh[J], t.bzl[J], t.bzh[J]);
                bp = sphx_within(d2, reach_b(J));
            }
            uint64_t bmask = __ballot(bp);
            while (bmask) {
                int J2 = j0 + __builtin_ctzll(bmask);
                bmask &= bmask - 1;
                body(J2);
            }
        }
    }
}

// for_candidates with the candidate loads run ahead: `load(J, beg, cnt)`
// issues the reads `body(J, v)` needs of bucket J (lane q: particle beg + q,
// q < cnt) and is called for the next surviving bucket before `body` runs on
// the current one, so a walk no longer waits on each bucket's loads.  The
// buckets' first particle and count come with the bucket culling, one bucket
// per lane, and reach the loads by readlane.  Buckets are met in the same
// order as by for_candidates.
template <class ReachS, class ReachB, class Load, class Body>
__device__ __forceinline__ void for_candidates_pf(const Tables &t, int r, const double *ib, int lane,
                                                  ReachS reach_s, ReachB reach_b, Load load,
                                                  Body body) {
    double axl = ib[0], axh = ib[1], ayl = ib[2], ayh = ib[3], azl = ib[4], azh = ib[5];
    int s0 = t.r_sbeg[r], s1 = t.r_send[r];
    decltype(load(0, 0, 0)) cur{};
    int Jc = -1;
    for (int sb = s0; sb < s1; sb += 64) {
        int S = sb + lane;
        bool sp = false;
        if (S < s1) {
            double d2 = sphx_box_d2(axl, axh, ayl, ayh, azl, azh, t.sxl[S], t.sxh[S], t.syl[S],
                                    t.syh[S], t.szl[S], t.szh[S]);
            sp = sphx_within(d2, reach_s(S));
        }
        uint64_t smask = __ballot(sp);
        while (smask) {
            int S2 = sb + __builtin_ctzll(smask);
            smask &= smask - 1;
            int j0 = t.s_bbeg[S2], j1 = t.s_bend[S2];
            int J = j0 + lane;
            bool bp = false;
            int vbeg = 0, vcnt = 0;
            if (J < j1) {
                double d2 = sphx_box_d2(axl, axh, ayl, ayh, azl, azh, t.bxl[J], t.bxh[J],
                                        t.byl[J], t.byh[J], t.bzl[J], t.bzh[J]);
                bp = sphx_within(d2, reach_b(J));
                vbeg = t.b_beg[J];
                vcnt = t.b_cnt[J];
            }
            uint64_t bmask = __ballot(bp);
            while (bmask) {
                int q = __builtin_ctzll(bmask);
                bmask &= bmask - 1;
                auto nxt = load(j0 + q, __builtin_amdgcn_readlane(vbeg, q),
                                __builtin_amdgcn_readlane(vcnt, q));
                if (Jc >= 0) body(Jc, cur);
                cur = nxt;
                Jc = j0 + q;
            }
        }
    }
    if (Jc >= 0) body(Jc, cur);
}

// Segments of a bucket.  64 consecutive Morton-sorted particles usually fill
// one compact cell, but where the curve jumps between distant cells their
// bounding box spans much of the run and would make every bucket of that
// region a candidate.  A bucket is therefore cut wherever two consecutive
// particles lie in different cells of Morton level t.r_lcut[run]; each
// segment walks the candidates of its own box and only its lanes accumulate.
// Candidates are still met in increasing order, so every sum is unchanged.
// Returns the ballot of lanes that start a segment (lane 0 never set).
__device__ __forceinline__ uint64_t bucket_cuts(const Tables &t, int r, int i, bool ok, int lane) {
    bool cut = false;
    if (t.keys && ok && lane > 0) {
        uint64_t x = t.keys[i] ^ t.keys[i - 1];
        if (x) {
            int hb = 63 - __builtin_clzll(x);
            int lev = t.mbits - 1 - hb / 3;  // Morton level of the cell boundary crossed
            cut = lev <= t.r_lcut[r];
        }
    }
    return __ballot(cut);
}

// End (exclusive) of the segment that starts at lane s0.
__device__ __forceinline__ int seg_end(uint64_t cuts, int s0, int cnt) {
    uint64_t rest = s0 >= 63 ? 0ull : cuts & (~0ull << (s0 + 1));
    int s1 = rest ? __builtin_ctzll(rest) : cnt;
    return s1 < cnt ? s1 : cnt;
}

// Bounding box and max h of the lanes with `mine` set.
__device__ __forceinline__ void seg_box(bool mine, double x, double y, double z, double h,
                                        double *box, double *hmax) {
    box[0] = wave_min_d(mine ? x : INFINITY);
    box[1] = wave_max_d(mine ? x : -INFINITY);
    box[2] = wave_min_d(mine ? y : INFINITY);
    box[3] = wave_max_d(mine ? y : -INFINITY);
    box[4] = wave_min_d(mine ? z : INFINITY);
    box[5] = wave_max_d(mine ? z : -INFINITY);
    *hmax = wave_max_d(mine ? h : 0.0);
}

// The particles a walk group takes: a whole bucket, or in runs too small to
// fill the device with one wave per bucket, half of one (lanes 0..31 take
// the bucket's lanes 32h .. 32h + 31; lanes 32..63 idle), so that twice the
// groups hide the walk's load and LDS latency.  A half walks the candidates
// of its own box, exactly as a segment does, so every particle still meets
// its candidates in increasing order and its sums and counts are unchanged.
struct WalkSlot {
    int b;      // bucket
    int cnt;    // particles of the group (lanes 0 .. cnt-1)
    int first;  // first particle
    int bl0;    // bucket lane of lane 0
    int h;      // half (0 for a whole bucket)
};
__device__ __forceinline__ WalkSlot walk_slot(const Tables &t, int halves) {
    const int g = xcd_block(blockIdx.x, gridDim.x);
    WalkSlot w;
    if (halves == 2) {
        w.b = t.boff + (g >> 1);
        w.h = g & 1;
        w.bl0 = 32 * w.h;
        int c = t.b_cnt[w.b] - w.bl0;
        w.cnt = c < 0 ? 0 : (c > 32 ? 32 : c);
    } else {
        w.b = t.boff + g;
        w.h = 0;
        w.bl0 = 0;
        w.cnt = t.b_cnt[w.b];
    }
    w.first = t.b_beg[w.b] + w.bl0;
    return w;
}

// Candidate filter: lane q holds particle q of a staged candidate bucket and
// keeps it when it may lie within `reach` of the target group's box (the same
// conservative FP64 test as the bucket culling).  A dropped particle is out of
// reach of every target particle, so the kept ones, compacted in their
// original order, give exactly the sums of the unfiltered loop.
struct Kept {
    int n, slot;
};
__device__ __forceinline__ Kept keep_mask(bool keep, int lane) {
    uint64_t m = __ballot(keep);
    Kept k;
    k.n = __popcll(m);
    k.slot = __popcll(m & ((1ull << lane) - 1ull));
    return k;
}

__device__ __forceinline__ bool near_box(const double *bx, double x, double y, double z,
                                         double reach) {
    return sphx_within(sphx_box_d2(bx[0], bx[1], bx[2], bx[3], bx[4], bx[5], x, x, y, y, z, z),
                       reach);
}

struct SmoothArgs {
    int32_t nngb, ndown, nup, max_iter;
    int32_t predictor;
    const RunParams *rp;
    int32_t *h_iters;  // per particle, nullable
    int64_t *counters;
    int32_t halves;    // 1: a group per bucket; 2: a group per half bucket
    double hstep;      // RULE 2: the step as a fraction of h_prev (2 x the reference's coeff)
    double hmin;       // RULE 5: the floor on h after eleven steps (0: none)
};

// RULE 0: the count-band iteration (a log-space Newton step, kept inside the
// bracket by bisection).  RULE 2: the reference's own rule (LIBSX3:947-1006,
// LIBV2:7-64; do_smoothingX4 with ten times the step): fixed steps of
// a.hstep h_prev (0.002 or 0.02) from 2 h_prev (or the eq. 31
// start, in the reference's order of operations), counts by sqrt(r^2) <= s as
// the reference tests, the largest radius taken after max_iter + 1 passes.
// RULE 4: the LIBV7 driver's rule (LIBV7:255-328): steps of 0.01 h_prev,
// halved on every jump across the band, the last radius after 501 passes,
// counts that stop at nup + 1.  RULE 5: h_smooth_fast (LIBSX:1133-1184):
// from h_prev with no predictor, h -+= 0.003 h on the current h, the count
// strict (sqrt(r^2) < 2h), the last h after SPHX_H5_PASSES passes.
template <int RULE>
__global__ __launch_bounds__(64) void k_smoothing(Tables t, Soa s, SmoothArgs a) {
    __shared__ double sx[64], sy[64], sz[64];
    const WalkSlot ws = walk_slot(t, a.halves);
    int b = ws.b;
    int lane = threadIdx.x;
    int cnt = ws.cnt;
    int i = ws.first + lane;
    bool ok = lane < cnt;
    int r = t.b_run[b];
    int64_t nrun = (int64_t)t.r_pend[r] - t.r_pbeg[r];
    double xi = ok ? s.x[i] : 0.0, yi = ok ? s.y[i] : 0.0, zi = ok ? s.z[i] : 0.0;
    double hprev = ok ? s.h[i] : 0.0;
    const double ib[6] = {t.bxl[b], t.bxh[b], t.byl[b], t.byh[b], t.bzl[b], t.bzh[b]};
    const uint64_t cuts = bucket_cuts(t, r, i, ok, lane);
    double sr;
    {
#pragma clang fp contract(off)
        sr = 2.0 * hprev;
        if (RULE != 5 && a.predictor && ok) {
            double dt = a.rp[r].dt;
            double f = sphx_h_factor(a.nngb, s.k[i]);
            double t1 = 0.5 * hprev;
            double t2 = 1.0 + f;
            double t3 = t1 * t2;
            double t6;
            if (RULE == 4) {         // LIBV7:265-266: ((1/3 h) divV) dt
                double u = (1.0 / 3.0) * hprev;
                u = u * s.divv[i];
                t6 = u * dt;
            } else if (RULE == 2) {  // LIBSX3:961: (1/3 (h divV)) dt
                double u = hprev * s.divv[i];
                u = (1.0 / 3.0) * u;
                t6 = u * dt;
            } else {
                double u = hprev * s.divv[i];
                u = u * dt;
                t6 = u / 3.0;
            }
            double sp = 2.0 * (t3 + t6);
            if (sp > 0.0 && sp < INFINITY) sr = sp;
        }
    }
    double lo = 0.0, hi = 0.0;
    int klo = 0, khi = 0, kfin = 0;
    bool have_lo = false, have_hi = false, done = !ok, conv = false;
    int it = 0;
    double coef = a.hstep;  // RULES 2, 4: the step over h_prev (RULE 4 halves it)
    double smax = sr;       // RULE 2: the largest radius tried
    int kpre = a.nup;       // RULE 4: the previous pass's count
    while (true) {
        bool active = ok && !done;
        if (!__any(active)) break;
        // once some lanes have settled, the walk culls against the box of the
        // lanes still counting: only their counts are read
        const bool all_active = !__any(ok && !active);
        double s2 = sphx_sq(sr);
        int kc = 0;
        for (int g0 = 0; g0 < cnt;) {
            int g1 = seg_end(cuts, g0, cnt);
            bool mine = lane >= g0 && lane < g1;
            double R = wave_max_d((active && mine) ? sr : 0.0);
            double gb[6];
            if (cuts || !all_active || a.halves == 2) {
                double hdummy;
                seg_box(active && mine, xi, yi, zi, 0.0, gb, &hdummy);
            } else {
                for (int k = 0; k < 6; ++k) gb[k] = ib[k];
            }
            g0 = g1;
            if (!(R > 0.0)) continue;  // no active lane in this segment
            struct Cand {
                double x, y, z;
                int n;
            };
            for_candidates_pf(
                t, r, gb, lane, [&](int) { return R; }, [&](int) { return R; },
                [&](int, int jb, int jc) {
                    Cand c{0.0, 0.0, 0.0, jc};
                    if (lane < jc) {
                        c.x = s.x[jb + lane];
                        c.y = s.y[jb + lane];
                        c.z = s.z[jb + lane];
                    }
                    return c;
                },
                [&](int, const Cand &c) {
                    double xj = c.x, yj = c.y, zj = c.z;
                    bool keep = false;
                    if (lane < c.n) keep = near_box(gb, xj, yj, zj, R);
                    Kept k = keep_mask(keep, lane);
                    if (k.n == 0) return;
                    __syncthreads();
                    if (keep) {
                        sx[k.slot] = xj;
                        sy[k.slot] = yj;
                        sz[k.slot] = zj;
                    }
                    __syncthreads();
                    // candidates four at a time, their LDS loads issued first
                    int q = 0;
                    for (; q + 4 <= k.n; q += 4) {
                        double cx[4], cy[4], cz[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            cx[u] = sx[q + u]; cy[u] = sy[q + u]; cz[u] = sz[q + u];
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            double r2 = sphx_r2(xi - cx[u], yi - cy[u], zi - cz[u]);
                            kc += (mine && (RULE == 5 ? sqrt(r2) < sr : (RULE >= 2 ? sqrt(r2) <= sr : r2 <= s2))) ? 1 : 0;
                        }
                    }
                    for (; q < k.n; ++q) {
                        double r2 = sphx_r2(xi - sx[q], yi - sy[q], zi - sz[q]);
                        kc += (mine && (RULE == 5 ? sqrt(r2) < sr : (RULE >= 2 ? sqrt(r2) <= sr : r2 <= s2))) ? 1 : 0;
                    }
                });
        }
        if (active) {
#pragma clang fp contract(off)
            ++it;
            if (RULE >= 2) {
                if (RULE == 4 && kc > a.nup) kc = a.nup + 1;  // LIBV7 stops counting there
                const bool in_band = kc >= a.ndown && kc <= a.nup;
                // LIBSX:1226-1228: after the eleventh step an h under the floor is
                // set to it and ends the iteration, in the band or not
                const bool floor_h = RULE == 5 && a.hmin > 0.0 && it > 11 && 0.5 * sr < a.hmin;
                if (RULE == 4 && ((kc < a.ndown && kpre > a.nup) || (kc > a.nup && kpre < a.ndown)))
                    coef = 0.5 * coef;  // a jump across the band
                if (RULE == 5) {
                    const double hc = 0.5 * sr;
                    if (kc < a.ndown) sr = 2.0 * (hc + 0.003 * hc);
                    if (kc > a.nup) sr = 2.0 * (hc - 0.003 * hc);
                } else {
                    if (kc < a.ndown) sr = sr + coef * hprev;
                    if (kc > a.nup) sr = sr - coef * hprev;
                }
                if (sr > smax) smax = sr;
                kpre = kc;
                kfin = kc;
                if (floor_h) {
                    sr = 2.0 * a.hmin;
                    kfin = kc;
                    done = true;
                    conv = true;
                } else if (in_band) {  // before the pass limit, as the reference tests
                    done = true;
                    conv = true;
                } else if (it > (RULE == 4 ? SPHX_H4_PASSES : (RULE == 5 ? SPHX_H5_PASSES : a.max_iter))) {
                    // the pass limit: RULE 2 takes the largest radius, RULES 4, 5 the last
                    if (RULE == 2) sr = smax;
                    done = true;
                }
            } else if (kc >= a.ndown && kc <= a.nup) {
                done = true;
                conv = true;
                kfin = kc;
            } else {
                if (kc < a.ndown) {
                    lo = sr; klo = kc; have_lo = true;
                    if ((int64_t)kc >= nrun) done = true;
                } else {
                    hi = sr; khi = kc; have_hi = true;
                }
                if (!done) {
                    double snew = sr * sphx_h_factor(a.nngb, kc);
                    if (have_lo && have_hi) {
                        if (!(snew > lo && snew < hi)) snew = 0.5 * (lo + hi);
                        if (!(snew > lo && snew < hi)) done = true;
                    }
                    if (!done) {
                        sr = snew;
                        if (it >= a.max_iter) done = true;
                    }
                }
            }
        }
    }
    if (ok) {
        if (!conv) {
            if (RULE == 0) {
                if (have_hi) { sr = hi; kfin = khi; }
                else { sr = lo; kfin = klo; }
            }
            atomicAdd((unsigned long long *)&a.counters[CNT_H_CAPPED], 1ull);
        }
        s.h[i] = 0.5 * sr;
        s.k[i] = kfin;
        if (a.h_iters) a.h_iters[i] = it;
    }
    int itmax = wave_max_i(ok ? it : 0);
    unsigned long long itsum = 0;
    {
        int v = ok ? it : 0;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        itsum = (unsigned long long)v;
    }
    if (lane == 0) {
        atomicMax((unsigned long long *)&a.counters[CNT_H_ITERS_MAX], (unsigned long long)itmax);
        atomicAdd((unsigned long long *)&a.counters[CNT_H_ITERS_SUM], itsum);
    }
}

// Cubic spline W / (sigma / h^3) and grad-W factor F / (sigma / h^5) in FP32.
__device__ __forceinline__ float w_shape(float q) {
    // both pieces, then a select: the lanes of a wave straddle q = 1, so a
    // branch would run both anyway and add the exec-mask bookkeeping
    float w1 = 1.0f - 1.5f * q * q + 0.75f * q * q * q;
    float w = fmaxf(2.0f - q, 0.0f);
    float w2 = 0.25f * w * w * w;
    return q <= 1.0f ? w1 : w2;
}
__device__ __forceinline__ float f_shape(float q) {
    if (q <= 1.0f) return -3.0f + 2.25f * q;
    float w = fmaxf(2.0f - q, 0.0f);
    return -0.75f * w * w * __builtin_amdgcn_rcpf(q);
}

struct PairGeom {
    float dx, dy, dz, r2, hinv, q;
};

__device__ __forceinline__ PairGeom pair_geom(double dx, double dy, double dz, double r2,
                                              double t) {
    PairGeom g;
    g.dx = (float)dx;
    g.dy = (float)dy;
    g.dz = (float)dz;
    g.r2 = (float)r2;
    float hij = 0.5f * (float)t;
    g.hinv = __builtin_amdgcn_rcpf(hij);
    float r = __builtin_amdgcn_sqrtf(g.r2);
    g.q = r * g.hinv;
    return g;
}

// The pair terms of the div/curl and hydro passes, shared by the candidate
// walks and the neighbour-list passes so that both evaluate every pair with
// the same instructions (the list passes are then bit-identical to the walks).
__device__ __forceinline__ void divcurl_pair(double vxi, double vyi, double vzi, double vxj,
                                             double vyj, double vzj, float mj, double dx,
                                             double dy, double dz, double r2, double tt,
                                             KahanF &kd, KahanF &kx, KahanF &ky, KahanF &kz) {
    PairGeom g = pair_geom(dx, dy, dz, r2, tt);
    float h5 = g.hinv * g.hinv;
    h5 = h5 * h5 * g.hinv;
    float F = SPHX_SIGMA_F * h5 * f_shape(g.q) * mj;
    float vx = (float)(vxi - vxj), vy = (float)(vyi - vyj), vz = (float)(vzi - vzj);
    kd.add(F * (vx * g.dx + vy * g.dy + vz * g.dz));
    kx.add(F * (vy * g.dz - vz * g.dy));
    ky.add(F * (vz * g.dx - vx * g.dz));
    kz.add(F * (vx * g.dy - vy * g.dx));
}

struct HydroArgs {
    float alpha, beta, eta2, w_eps;
    int32_t av_form, balsara;
};

// Particle i's side of the hydro pair terms.
struct HydroI {
    double vx, vy, vz;
    float rho, pr, c, f;  // rho, P / rho^2, sound speed, Balsara factor
};

__device__ __forceinline__ void hydro_pair(const HydroArgs &a, const HydroI &I, double vxj,
                                           double vyj, double vzj, float mj, float rhoj,
                                           float prj, float cj, float fj, double dx, double dy,
                                           double dz, double r2, double tt, KahanF &kx,
                                           KahanF &ky, KahanF &kz, KahanF &kdu) {
    PairGeom g = pair_geom(dx, dy, dz, r2, tt);
    float h5 = g.hinv * g.hinv;
    h5 = h5 * h5 * g.hinv;
    float F = SPHX_SIGMA_F * h5 * f_shape(g.q);
    float vx = (float)(I.vx - vxj), vy = (float)(I.vy - vyj), vz = (float)(I.vz - vzj);
    float vr = vx * g.dx + vy * g.dy + vz * g.dz;
    float pij = 0.0f;
    if (vr <= 0.0f) {
        float rhob = 0.5f * (I.rho + rhoj);
        float hij = 0.5f * (float)tt;
        if (a.av_form == 0) {
            float mu = hij * vr / (g.r2 + hij * hij * a.eta2);
            pij = (-a.alpha * 0.5f * (I.c + cj) * mu + a.beta * mu * mu) / rhob;
        } else {
            float w = vr / (__builtin_amdgcn_sqrtf(g.r2) + a.w_eps);
            pij = -0.5f * a.alpha * (I.c + cj - 3.0f * w) * w / rhob;
        }
        pij *= 0.5f * (I.f + fj);
    }
    float mF = mj * F;
    float sc_ = mF * (I.pr + prj + pij);
    kx.add(-sc_ * g.dx);
    ky.add(-sc_ * g.dy);
    kz.add(-sc_ * g.dz);
    kdu.add(mF * (I.pr + 0.5f * pij) * vr);
}

__device__ __forceinline__ double balsara_f(double dv, double cv, double c, double h) {
    double a = fabs(dv);
    return a / (a + cv + 0.0001 * c / h);
}

// Position of entry k of lane L in its bucket's list block: lanes hold their
// entries in runs of 8 (32 bytes), the runs of the 64 lanes side by side, so
// that a lane fills its own 32-byte piece as it accepts neighbours (four lanes
// complete a 128-byte line together) and iteration k of a list pass reads
// 64 runs that iterations k + 1 .. k + 7 read again.
__device__ __forceinline__ int64_t nbr_slot(int k, int lane) {
    return ((int64_t)(k >> 3) * 64 + lane) * 8 + (k & 7);
}

// Calls f(j) for entries 0 .. len-1 of this lane's list, every lane on its
// own entry k at once (maxlen: the wave's longest list); each 8-entry run
// comes in with two 16-byte loads.
template <class F>
__device__ __forceinline__ void for_list(const int32_t *lp, int lane, int len, int maxlen, F &&f) {
    for (int k0 = 0; k0 < maxlen; k0 += 8) {
        int4 a = make_int4(0, 0, 0, 0), c = a;
        if (k0 < len) {
            const int4 *p = reinterpret_cast<const int4 *>(lp + nbr_slot(k0, lane));
            a = p[0];
            c = p[1];
        }
#pragma unroll 1
        for (int q = 0; q < 8 && k0 + q < maxlen; ++q) {
            if (k0 + q < len) f(a.x);
            a.x = a.y; a.y = a.z; a.z = a.w; a.w = c.x;
            c.x = c.y; c.y = c.z; c.z = c.w;
        }
    }
}

struct DensArgs {
    int32_t halves;       // 1: a group per bucket; 2: a group per half bucket
    double gamma;
    int32_t *nsym;        // symmetric neighbour count, nullable
    int64_t *counters;
    // neighbour lists for the div/curl and hydro passes (all null: none written)
    int32_t *lst;         // [bucket][entry < SPHX_NBR_MAX][lane]: accepted j in walk order
    PairRec *rec;         // per particle: the j-side values of the pair passes
    int32_t *ovf;         // two per bucket (halves): 1 when a lane's list did not fit
};

__global__ __launch_bounds__(64) void k_density(Tables t, Soa s, DensArgs a) {
    __shared__ double sx[64], sy[64], sz[64], sh[64];
    __shared__ float sm[64];
    __shared__ int32_t sj[64];
    const WalkSlot ws = walk_slot(t, a.halves);
    int b = ws.b;
    int lane = threadIdx.x;
    int cnt = ws.cnt;
    int i = ws.first + lane;
    bool ok = lane < cnt;
    double xi = ok ? s.x[i] : 0.0, yi = ok ? s.y[i] : 0.0, zi = ok ? s.z[i] : 0.0;
    double hi = ok ? s.h[i] : 0.0;
    double hI = t.bhmax[b];
    const double ib[6] = {t.bxl[b], t.bxh[b], t.byl[b], t.byh[b], t.bzl[b], t.bzh[b]};
    KahanF acc;
    acc.init();
    int nn = 0;
    int ncand = 0;
    int32_t *lst = a.lst ? a.lst + (int64_t)b * SPHX_NBR_MAX * 64 : nullptr;
    // this lane's first run of the block: entry k at ((k / 8) * 64 * 8) + k mod 8
    // from here (nbr_slot(k, L) - nbr_slot(0, L), L its lane in the bucket)
    int32_t *lane_lst = lst ? lst + (ws.bl0 + lane) * 8 : nullptr;
    const int r = t.b_run[b];
    const uint64_t cuts = bucket_cuts(t, r, i, ok, lane);
    for (int g0 = 0; g0 < cnt;) {
      const int g1 = seg_end(cuts, g0, cnt);
      const bool mine = lane >= g0 && lane < g1;
      double gb[6], hS;
      if (cuts || a.halves == 2) {
          seg_box(ok && mine, xi, yi, zi, hi, gb, &hS);
      } else {
          for (int k = 0; k < 6; ++k) gb[k] = ib[k];
          hS = hI;
      }
      g0 = g1;
      struct Cand {
          double x, y, z, h, m;
          int beg, n;
      };
      for_candidates_pf(
        t, r, gb, lane, [&](int S) { return hS + t.shmax[S]; }, [&](int J) { return hS + t.bhmax[J]; },
        [&](int, int jb, int jc) {
            Cand c{0.0, 0.0, 0.0, 0.0, 0.0, jb, jc};
            if (lane < jc) {
                c.x = s.x[jb + lane];
                c.y = s.y[jb + lane];
                c.z = s.z[jb + lane];
                c.h = s.h[jb + lane];
                c.m = s.m[jb + lane];
            }
            return c;
        },
        [&](int, const Cand &c) {
            ++ncand;
            double xj = c.x, yj = c.y, zj = c.z, hj = c.h;
            bool keep = false;
            if (lane < c.n) keep = near_box(gb, xj, yj, zj, hS + hj);
            Kept k = keep_mask(keep, lane);
            if (k.n == 0) return;
            __syncthreads();
            if (keep) {
                sx[k.slot] = xj;
                sy[k.slot] = yj;
                sz[k.slot] = zj;
                sh[k.slot] = hj;
                sm[k.slot] = (float)c.m;
                sj[k.slot] = c.beg + lane;
            }
            __syncthreads();
            // candidates four at a time, their LDS loads issued before the tests
            // (each pair is met in the same order as one at a time)
            auto pair = [&](int q, double dx, double dy, double dz, double r2, double tt) {
                    PairGeom g = pair_geom(dx, dy, dz, r2, tt);
                    float h3 = g.hinv * g.hinv * g.hinv;
                    acc.add(sm[q] * (SPHX_SIGMA_F * h3 * w_shape(g.q)));
                    if (lst && nn < SPHX_NBR_MAX) lane_lst[((nn >> 3) << 9) | (nn & 7)] = sj[q];
                    ++nn;
                };
            int q0 = 0;
            for (; q0 + 4 <= k.n; q0 += 4) {
                double cx[4], cy[4], cz[4], ch[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    cx[u] = sx[q0 + u]; cy[u] = sy[q0 + u]; cz[u] = sz[q0 + u]; ch[u] = sh[q0 + u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    double dx = xi - cx[u], dy = yi - cy[u], dz = zi - cz[u];
                    double r2 = sphx_r2(dx, dy, dz);
                    double tt = hi + ch[u];
                    if (mine && r2 <= sphx_sq(tt)) pair(q0 + u, dx, dy, dz, r2, tt);
                }
            }
            for (int q = q0; q < k.n; ++q) {
                double dx = xi - sx[q], dy = yi - sy[q], dz = zi - sz[q];
                double r2 = sphx_r2(dx, dy, dz);
                double tt = hi + sh[q];
                if (mine && r2 <= sphx_sq(tt)) pair(q, dx, dy, dz, r2, tt);
            }
        });
    }
    if (ok) {
        double rho = acc.value();
        double u = s.u[i];
        double P = (a.gamma - 1.0) * rho * u;
        double c = sqrt(a.gamma * (a.gamma - 1.0) * u);
        s.rho[i] = rho;
        s.P[i] = P;
        s.c[i] = c;
        if (a.nsym) a.nsym[i] = nn;
        if (a.rec) {
            // the values the pair passes stage for a neighbour j, as they stage them
            PairRec q;
            q.x = xi; q.y = yi; q.z = zi; q.h = hi;
            q.vx = s.vx[i]; q.vy = s.vy[i]; q.vz = s.vz[i];
            q.m = (float)s.m[i];
            q.rho = (float)rho;
            q.pr = (float)(P / (rho * rho));
            q.c = (float)c;
            q.len = nn;
            q.pad = 0;
            a.rec[i] = q;
        }
    }
    if (a.ovf) {
        // two flags per bucket, one per half (a whole bucket sets both)
        uint64_t o = __ballot(ok && nn > SPHX_NBR_MAX);
        if (lane == 0) {
            a.ovf[2 * b + ws.h] = o ? 1 : 0;
            if (a.halves == 1) a.ovf[2 * b + 1] = 0;
            if (o && a.counters) atomicAdd((unsigned long long *)&a.counters[CNT_NBR_OVF], 1ull);
        }
    }
    if (lane == 0 && a.counters)
        atomicAdd((unsigned long long *)&a.counters[CNT_CAND], (unsigned long long)ncand);
}

__device__ __forceinline__ void divcurl_store(const Soa &s, int i, bool ok, const KahanF &kd,
                                              const KahanF &kx, const KahanF &ky,
                                              const KahanF &kz) {
    if (ok) {
        double ir = 1.0 / s.rho[i];
        s.divv[i] = -kd.value() * ir;
        double cx = kx.value() * ir, cy = ky.value() * ir, cz = kz.value() * ir;
        s.curlv[i] = sqrt(cx * cx + cy * cy + cz * cz);
    }
}

// div/curl of the particles of bucket b by the candidate walk.
__device__ __forceinline__ void divcurl_bucket(const Tables &t, const Soa &s, int b) {
    __shared__ double sx[64], sy[64], sz[64], sh[64], svx[64], svy[64], svz[64];
    __shared__ float sm[64];
    int lane = threadIdx.x;
    int cnt = t.b_cnt[b];
    int i = t.b_beg[b] + lane;
    bool ok = lane < cnt;
    double xi = ok ? s.x[i] : 0.0, yi = ok ? s.y[i] : 0.0, zi = ok ? s.z[i] : 0.0;
    double hi = ok ? s.h[i] : 0.0;
    double vxi = ok ? s.vx[i] : 0.0, vyi = ok ? s.vy[i] : 0.0, vzi = ok ? s.vz[i] : 0.0;
    double hI = t.bhmax[b];
    const double ib[6] = {t.bxl[b], t.bxh[b], t.byl[b], t.byh[b], t.bzl[b], t.bzh[b]};
    KahanF kd, kx, ky, kz;
    kd.init(); kx.init(); ky.init(); kz.init();
    const int r = t.b_run[b];
    const uint64_t cuts = bucket_cuts(t, r, i, ok, lane);
    for (int g0 = 0; g0 < cnt;) {
      const int g1 = seg_end(cuts, g0, cnt);
      const bool mine = lane >= g0 && lane < g1;
      double gb[6], hS;
      if (cuts) {
          seg_box(ok && mine, xi, yi, zi, hi, gb, &hS);
      } else {
          for (int k = 0; k < 6; ++k) gb[k] = ib[k];
          hS = hI;
      }
      g0 = g1;
      for_candidates(
        t, r, gb, lane, [&](int S) { return hS + t.shmax[S]; }, [&](int J) { return hS + t.bhmax[J]; },
        [&](int J) {
            int jc = t.b_cnt[J], j0 = t.b_beg[J];
            int j = j0 + lane;
            double xj = 0.0, yj = 0.0, zj = 0.0, hj = 0.0;
            bool keep = false;
            if (lane < jc) {
                xj = s.x[j]; yj = s.y[j]; zj = s.z[j]; hj = s.h[j];
                keep = near_box(gb, xj, yj, zj, hS + hj);
            }
            Kept k = keep_mask(keep, lane);
            if (k.n == 0) return;
            __syncthreads();
            if (keep) {
                int q = k.slot;
                sx[q] = xj; sy[q] = yj; sz[q] = zj; sh[q] = hj;
                svx[q] = s.vx[j]; svy[q] = s.vy[j]; svz[q] = s.vz[j];
                sm[q] = (float)s.m[j];
            }
            __syncthreads();
            // candidates four at a time, their LDS loads issued before the tests
            // (each pair is met in the same order as one at a time)
            auto pair = [&](int q, double dx, double dy, double dz, double r2, double tt) {
                    divcurl_pair(vxi, vyi, vzi, svx[q], svy[q], svz[q], sm[q], dx, dy, dz, r2,
                                 tt, kd, kx, ky, kz);
                };
            int q0 = 0;
            for (; q0 + 4 <= k.n; q0 += 4) {
                double cx[4], cy[4], cz[4], ch[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    cx[u] = sx[q0 + u]; cy[u] = sy[q0 + u]; cz[u] = sz[q0 + u]; ch[u] = sh[q0 + u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    double dx = xi - cx[u], dy = yi - cy[u], dz = zi - cz[u];
                    double r2 = sphx_r2(dx, dy, dz);
                    double tt = hi + ch[u];
                    if (mine && r2 <= sphx_sq(tt)) pair(q0 + u, dx, dy, dz, r2, tt);
                }
            }
            for (int q = q0; q < k.n; ++q) {
                double dx = xi - sx[q], dy = yi - sy[q], dz = zi - sz[q];
                double r2 = sphx_r2(dx, dy, dz);
                double tt = hi + sh[q];
                if (mine && r2 <= sphx_sq(tt)) pair(q, dx, dy, dz, r2, tt);
            }
        });
    }
    divcurl_store(s, i, ok, kd, kx, ky, kz);
}

__global__ __launch_bounds__(64) void k_divcurl(Tables t, Soa s) {
    divcurl_bucket(t, s, t.boff + xcd_block(blockIdx.x, gridDim.x));
}

// Hydro acceleration and du/dt of the particles of bucket b by the candidate walk.
__device__ __forceinline__ void hydro_bucket(const Tables &t, const Soa &s, const HydroArgs &a,
                                             int b, bool add_g = false) {
    __shared__ double sx[64], sy[64], sz[64], sh[64], svx[64], svy[64], svz[64];
    __shared__ float sm[64], srho[64], spr[64], sc[64], sf[64];
    int lane = threadIdx.x;
    int cnt = t.b_cnt[b];
    int i = t.b_beg[b] + lane;
    bool ok = lane < cnt;
    double xi = ok ? s.x[i] : 0.0, yi = ok ? s.y[i] : 0.0, zi = ok ? s.z[i] : 0.0;
    double hi = ok ? s.h[i] : 0.0;
    HydroI I;
    I.vx = ok ? s.vx[i] : 0.0;
    I.vy = ok ? s.vy[i] : 0.0;
    I.vz = ok ? s.vz[i] : 0.0;
    I.rho = ok ? (float)s.rho[i] : 1.0f;
    I.pr = ok ? (float)(s.P[i] / (s.rho[i] * s.rho[i])) : 0.0f;
    I.c = ok ? (float)s.c[i] : 0.0f;
    I.f = (ok && a.balsara) ? (float)balsara_f(s.divv[i], s.curlv[i], s.c[i], s.h[i]) : 1.0f;
    double hI = t.bhmax[b];
    const double ib[6] = {t.bxl[b], t.bxh[b], t.byl[b], t.byh[b], t.bzl[b], t.bzh[b]};
    KahanF kx, ky, kz, kdu;
    kx.init(); ky.init(); kz.init(); kdu.init();
    const int r = t.b_run[b];
    const uint64_t cuts = bucket_cuts(t, r, i, ok, lane);
    for (int g0 = 0; g0 < cnt;) {
      const int g1 = seg_end(cuts, g0, cnt);
      const bool mine = lane >= g0 && lane < g1;
      double gb[6], hS;
      if (cuts) {
          seg_box(ok && mine, xi, yi, zi, hi, gb, &hS);
      } else {
          for (int k = 0; k < 6; ++k) gb[k] = ib[k];
          hS = hI;
      }
      g0 = g1;
      for_candidates(
        t, r, gb, lane, [&](int S) { return hS + t.shmax[S]; }, [&](int J) { return hS + t.bhmax[J]; },
        [&](int J) {
            int jc = t.b_cnt[J], j0 = t.b_beg[J];
            int j = j0 + lane;
            double xj = 0.0, yj = 0.0, zj = 0.0, hj = 0.0;
            bool keep = false;
            if (lane < jc) {
                xj = s.x[j]; yj = s.y[j]; zj = s.z[j]; hj = s.h[j];
                keep = near_box(gb, xj, yj, zj, hS + hj);
            }
            Kept k = keep_mask(keep, lane);
            if (k.n == 0) return;
            __syncthreads();
            if (keep) {
                int q = k.slot;
                sx[q] = xj; sy[q] = yj; sz[q] = zj; sh[q] = hj;
                svx[q] = s.vx[j]; svy[q] = s.vy[j]; svz[q] = s.vz[j];
                double rj = s.rho[j];
                sm[q] = (float)s.m[j];
                srho[q] = (float)rj;
                spr[q] = (float)(s.P[j] / (rj * rj));
                sc[q] = (float)s.c[j];
                sf[q] = a.balsara ? (float)balsara_f(s.divv[j], s.curlv[j], s.c[j], s.h[j]) : 1.0f;
            }
            __syncthreads();
            // candidates four at a time, their LDS loads issued before the tests
            // (each pair is met in the same order as one at a time)
            auto pair = [&](int q, double dx, double dy, double dz, double r2, double tt) {
                    hydro_pair(a, I, svx[q], svy[q], svz[q], sm[q], srho[q], spr[q], sc[q],
                               sf[q], dx, dy, dz, r2, tt, kx, ky, kz, kdu);
                };
            int q0 = 0;
            for (; q0 + 4 <= k.n; q0 += 4) {
                double cx[4], cy[4], cz[4], ch[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    cx[u] = sx[q0 + u]; cy[u] = sy[q0 + u]; cz[u] = sz[q0 + u]; ch[u] = sh[q0 + u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    double dx = xi - cx[u], dy = yi - cy[u], dz = zi - cz[u];
                    double r2 = sphx_r2(dx, dy, dz);
                    double tt = hi + ch[u];
                    if (mine && r2 <= sphx_sq(tt)) pair(q0 + u, dx, dy, dz, r2, tt);
                }
            }
            for (int q = q0; q < k.n; ++q) {
                double dx = xi - sx[q], dy = yi - sy[q], dz = zi - sz[q];
                double r2 = sphx_r2(dx, dy, dz);
                double tt = hi + sh[q];
                if (mine && r2 <= sphx_sq(tt)) pair(q, dx, dy, dz, r2, tt);
            }
        });
    }
    if (ok) {
        // add_g: the total acceleration, gravity + a_sph, as sphx_combine adds it
        s.ax[i] = add_g ? s.gx[i] + kx.value() : kx.value();
        s.ay[i] = add_g ? s.gy[i] + ky.value() : ky.value();
        s.az[i] = add_g ? s.gz[i] + kz.value() : kz.value();
        s.du[i] = kdu.value();
    }
}

__global__ __launch_bounds__(64) void k_hydro(Tables t, Soa s, HydroArgs a) {
    hydro_bucket(t, s, a, t.boff + xcd_block(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------------------
// Neighbour-list passes.  The density walk records, for every particle, the
// indices of the particles it accepted (r_ij^2 <= (h_i + h_j)^2) in the order
// it met them, and a record of the values the pair passes stage for a
// neighbour.  Positions and h do not change between the density pass and the
// hydro pass of a step, and div/curl and hydro accept exactly the pairs density
// accepts, so iterating a particle's list visits the same pairs in the same
// order as the candidate walk: with the shared pair terms the sums are
// bit-identical, without the walk's candidate tests or its wave-wide pair
// evaluations for candidates that only some lanes accept.  Buckets where a list
// overflowed (more than SPHX_NBR_MAX neighbours) are left to the walk.
// ---------------------------------------------------------------------------
struct ListArgs {
    const int32_t *lst;
    const PairRec *rec;
    const int32_t *ovf;
    const float *fbal;   // Balsara factor per particle (hydro with the switch on)
};

__device__ __forceinline__ double rec_dbl(uint32_t lo, uint32_t hi) {
    return __hiloint2double((int)hi, (int)lo);
}

// The position, h, velocity and mass of record j (four 16-byte loads).
struct RecA {
    double x, y, z, h, vx, vy, vz;
    float m, rho;
};
__device__ __forceinline__ RecA load_rec_a(const PairRec *rec, int j) {
    const uint4 *q = reinterpret_cast<const uint4 *>(rec + j);
    uint4 a = q[0], b = q[1], c = q[2], d = q[3];
    RecA o;
    o.x = rec_dbl(a.x, a.y); o.y = rec_dbl(a.z, a.w);
    o.z = rec_dbl(b.x, b.y); o.h = rec_dbl(b.z, b.w);
    o.vx = rec_dbl(c.x, c.y); o.vy = rec_dbl(c.z, c.w);
    o.vz = rec_dbl(d.x, d.y);
    o.m = __uint_as_float(d.z); o.rho = __uint_as_float(d.w);
    return o;
}
// P / rho^2, sound speed and list length of record j (the fifth 16 bytes).
__device__ __forceinline__ uint4 load_rec_b(const PairRec *rec, int j) {
    return reinterpret_cast<const uint4 *>(rec + j)[4];
}

__global__ __launch_bounds__(64) void k_divcurl_list(Tables t, Soa s, ListArgs L) {
    const int b = t.boff + xcd_block(blockIdx.x, gridDim.x);
    if (L.ovf[2 * b] | L.ovf[2 * b + 1]) return;  // the walk launch takes this bucket
    const int lane = threadIdx.x;
    const int cnt = t.b_cnt[b];
    const int i = t.b_beg[b] + lane;
    const bool ok = lane < cnt;
    RecA me{};
    int len = 0;
    if (ok) {
        me = load_rec_a(L.rec, i);
        len = (int)load_rec_b(L.rec, i).z;
    }
    const int maxlen = wave_max_i(len);
    const int32_t *lp = L.lst + (int64_t)b * SPHX_NBR_MAX * 64;
    KahanF kd, kx, ky, kz;
    kd.init(); kx.init(); ky.init(); kz.init();
    for_list(lp, lane, len, maxlen, [&](int j) {
        const RecA o = load_rec_a(L.rec, j);
        double dx = me.x - o.x, dy = me.y - o.y, dz = me.z - o.z;
        double r2 = sphx_r2(dx, dy, dz);
        double tt = me.h + o.h;
        divcurl_pair(me.vx, me.vy, me.vz, o.vx, o.vy, o.vz, o.m, dx, dy, dz, r2, tt, kd, kx, ky,
                     kz);
    });
    divcurl_store(s, i, ok, kd, kx, ky, kz);
}

__global__ __launch_bounds__(64) void k_hydro_list(Tables t, Soa s, HydroArgs a, ListArgs L) {
    const int b = t.boff + xcd_block(blockIdx.x, gridDim.x);
    if (L.ovf[2 * b] | L.ovf[2 * b + 1]) return;  // the walk launch takes this bucket
    const int lane = threadIdx.x;
    const int cnt = t.b_cnt[b];
    const int i = t.b_beg[b] + lane;
    const bool ok = lane < cnt;
    RecA me{};
    HydroI I;
    I.vx = I.vy = I.vz = 0.0;
    I.rho = 1.0f; I.pr = 0.0f; I.c = 0.0f; I.f = 1.0f;
    int len = 0;
    if (ok) {
        me = load_rec_a(L.rec, i);
        uint4 e = load_rec_b(L.rec, i);
        I.vx = me.vx; I.vy = me.vy; I.vz = me.vz;
        I.rho = me.rho;
        I.pr = __uint_as_float(e.x);
        I.c = __uint_as_float(e.y);
        if (a.balsara) I.f = L.fbal[i];
        len = (int)e.z;
    }
    const int maxlen = wave_max_i(len);
    const int32_t *lp = L.lst + (int64_t)b * SPHX_NBR_MAX * 64;
    KahanF kx, ky, kz, kdu;
    kx.init(); ky.init(); kz.init(); kdu.init();
    for_list(lp, lane, len, maxlen, [&](int j) {
        const RecA o = load_rec_a(L.rec, j);
        const uint4 e = load_rec_b(L.rec, j);
        const float fj = a.balsara ? L.fbal[j] : 1.0f;
        double dx = me.x - o.x, dy = me.y - o.y, dz = me.z - o.z;
        double r2 = sphx_r2(dx, dy, dz);
        double tt = me.h + o.h;
        hydro_pair(a, I, o.vx, o.vy, o.vz, o.m, o.rho, __uint_as_float(e.x), __uint_as_float(e.y),
                   fj, dx, dy, dz, r2, tt, kx, ky, kz, kdu);
    });
    if (ok) {
        s.ax[i] = kx.value();
        s.ay[i] = ky.value();
        s.az[i] = kz.value();
        s.du[i] = kdu.value();
    }
}

// div/curl and hydro of bucket b in one list pass, for steps without the
// Balsara switch (hydro then reads nothing div/curl produces): each
// neighbour's record is gathered once for both, and each of the eight sums
// meets its pairs in the list order, as the two separate passes do, so the
// results are the same.  The stored acceleration is the total, gravity +
// a_sph, added as sphx_combine adds it.
__device__ __forceinline__ void dchydro_bucket(const Tables &t, const Soa &s, const HydroArgs &a,
                                               const ListArgs &L, int b) {
    if (L.ovf[2 * b] | L.ovf[2 * b + 1]) return;  // the walk launches take this bucket
    const int lane = threadIdx.x;
    const int cnt = t.b_cnt[b];
    const int i = t.b_beg[b] + lane;
    const bool ok = lane < cnt;
    RecA me{};
    HydroI I;
    I.vx = I.vy = I.vz = 0.0;
    I.rho = 1.0f; I.pr = 0.0f; I.c = 0.0f; I.f = 1.0f;
    int len = 0;
    if (ok) {
        me = load_rec_a(L.rec, i);
        uint4 e = load_rec_b(L.rec, i);
        I.vx = me.vx; I.vy = me.vy; I.vz = me.vz;
        I.rho = me.rho;
        I.pr = __uint_as_float(e.x);
        I.c = __uint_as_float(e.y);
        len = (int)e.z;
    }
    const int maxlen = wave_max_i(len);
    const int32_t *lp = L.lst + (int64_t)b * SPHX_NBR_MAX * 64;
    KahanF kd, kx, ky, kz, hx, hy, hz, hdu;
    kd.init(); kx.init(); ky.init(); kz.init();
    hx.init(); hy.init(); hz.init(); hdu.init();
    for_list(lp, lane, len, maxlen, [&](int j) {
        const RecA o = load_rec_a(L.rec, j);
        const uint4 e = load_rec_b(L.rec, j);
        double dx = me.x - o.x, dy = me.y - o.y, dz = me.z - o.z;
        double r2 = sphx_r2(dx, dy, dz);
        double tt = me.h + o.h;
        divcurl_pair(me.vx, me.vy, me.vz, o.vx, o.vy, o.vz, o.m, dx, dy, dz, r2, tt, kd, kx, ky,
                     kz);
        hydro_pair(a, I, o.vx, o.vy, o.vz, o.m, o.rho, __uint_as_float(e.x), __uint_as_float(e.y),
                   1.0f, dx, dy, dz, r2, tt, hx, hy, hz, hdu);
    });
    divcurl_store(s, i, ok, kd, kx, ky, kz);
    if (ok) {
        s.ax[i] = s.gx[i] + hx.value();
        s.ay[i] = s.gy[i] + hy.value();
        s.az[i] = s.gz[i] + hz.value();
        s.du[i] = hdu.value();
    }
}

__global__ __launch_bounds__(64) void k_dchydro_list(Tables t, Soa s, HydroArgs a, ListArgs L) {
    dchydro_bucket(t, s, a, L, t.boff + xcd_block(blockIdx.x, gridDim.x));
}

// The walk for the buckets whose lists overflowed: each wave scans 64 flags
// at a time and walks the flagged buckets of its share.
__global__ __launch_bounds__(64) void k_divcurl_ovf(Tables t, Soa s, const int32_t *ovf, int nb) {
    const int lane = threadIdx.x;
    for (int base = blockIdx.x * 64; base < nb; base += gridDim.x * 64) {
        const int64_t fb = 2 * ((int64_t)t.boff + base + lane);
        uint64_t m = __ballot(base + lane < nb && (ovf[fb] | ovf[fb + 1]) != 0);
        while (m) {
            int q = __builtin_ctzll(m);
            m &= m - 1;
            divcurl_bucket(t, s, t.boff + base + q);
        }
    }
}

__global__ __launch_bounds__(64) void k_hydro_ovf(Tables t, Soa s, HydroArgs a,
                                                  const int32_t *ovf, int nb, int add_g) {
    const int lane = threadIdx.x;
    for (int base = blockIdx.x * 64; base < nb; base += gridDim.x * 64) {
        const int64_t fb = 2 * ((int64_t)t.boff + base + lane);
        uint64_t m = __ballot(base + lane < nb && (ovf[fb] | ovf[fb + 1]) != 0);
        while (m) {
            int q = __builtin_ctzll(m);
            m &= m - 1;
            hydro_bucket(t, s, a, t.boff + base + q, add_g != 0);
        }
    }
}

// Balsara factor of every particle of [p0, p1), as the walk computes it.
__global__ void k_fbal(Soa s, int64_t p0, int64_t p1, float *__restrict__ fbal) {
    int64_t i = p0 + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < p1) fbal[i] = (float)balsara_f(s.divv[i], s.curlv[i], s.c[i], s.h[i]);
}

// kNN smoothing length (do_smoothingX, LIBSX:762-783): h = 0.5 sqrt(v) with v
// the kth smallest r^2 (0-based, self included) from particle i to all
// particles of its run.  Each lane keeps a bracket: lo2 (exclusive, klo
// values <= lo2), up2 (a value with >= kth + 1 values <= it, once known) and
// the probe of the next walk.  A walk counts the values in (lo2, probe] and
// keeps up to KNN_CAP of them in LDS; when the rank falls inside and they
// fit, the value of that rank is selected exactly; otherwise the bracket
// grows (too few) or is bisected in value (too many).  The first probe is the
// radius of a count iteration whose band is [kth + 1, kth + 56], so one walk
// usually suffices.
constexpr int KNN_CAP = 128;

struct KnnArgs {
    int32_t kth;
    int32_t max_walks;
    int64_t *counters;  // lanes left unresolved -> CNT_H_CAPPED
};

__global__ __launch_bounds__(64) void k_knn(Tables t, Soa s, KnnArgs a) {
    __shared__ double sx[64], sy[64], sz[64];
    __shared__ double lst[KNN_CAP * 64];
    int b = t.boff + xcd_block(blockIdx.x, gridDim.x);
    int lane = threadIdx.x;
    int cnt = t.b_cnt[b];
    int i = t.b_beg[b] + lane;
    bool ok = lane < cnt;
    double xi = ok ? s.x[i] : 0.0, yi = ok ? s.y[i] : 0.0, zi = ok ? s.z[i] : 0.0;
    const int r = t.b_run[b];
    const uint64_t cuts = bucket_cuts(t, r, i, ok, lane);
    double lo2 = -1.0, up2 = INFINITY;
    double probe = ok ? sphx_sq(2.0 * s.h[i]) : 0.0;
    int klo = 0;
    bool done = !ok;
    double res = 0.0;
    for (int walk = 0; walk < a.max_walks && __any(!done); ++walk) {
        int cin = 0;
        for (int g0 = 0; g0 < cnt;) {
            const int g1 = seg_end(cuts, g0, cnt);
            const bool mine = lane >= g0 && lane < g1;
            const bool act = mine && !done;
            g0 = g1;
            // a probe of 0 is legitimate (coincident particles): walk whenever
            // a lane of the segment is active; reach 0 still finds the
            // buckets whose box contains the particle
            if (!__any(act)) continue;
            double P = wave_max_d(act ? probe : 0.0);
            double R = sqrt(P);
            double gb[6], hdummy;
            seg_box(ok && mine, xi, yi, zi, 0.0, gb, &hdummy);
            for_candidates(
                t, r, gb, lane, [&](int) { return R; }, [&](int) { return R; },
                [&](int J) {
                    int jc = t.b_cnt[J], j0 = t.b_beg[J];
                    double xj = 0.0, yj = 0.0, zj = 0.0;
                    bool keep = false;
                    if (lane < jc) {
                        xj = s.x[j0 + lane];
                        yj = s.y[j0 + lane];
                        zj = s.z[j0 + lane];
                        keep = near_box(gb, xj, yj, zj, R);
                    }
                    Kept k = keep_mask(keep, lane);
                    if (k.n == 0) return;
                    __syncthreads();
                    if (keep) {
                        sx[k.slot] = xj;
                        sy[k.slot] = yj;
                        sz[k.slot] = zj;
                    }
                    __syncthreads();
                    // candidates four at a time, their LDS loads issued first (the
                    // values are appended in the same order)
                    int q = 0;
                    for (; q + 4 <= k.n; q += 4) {
                        double cx[4], cy[4], cz[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            cx[u] = sx[q + u]; cy[u] = sy[q + u]; cz[u] = sz[q + u];
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            double v = sphx_r2(cx[u] - xi, cy[u] - yi, cz[u] - zi);
                            if (act && v > lo2 && v <= probe) {
                                if (cin < KNN_CAP) lst[cin * 64 + lane] = v;
                                ++cin;
                            }
                        }
                    }
                    for (; q < k.n; ++q) {
                        double v = sphx_r2(sx[q] - xi, sy[q] - yi, sz[q] - zi);
                        if (act && v > lo2 && v <= probe) {
                            if (cin < KNN_CAP) lst[cin * 64 + lane] = v;
                            ++cin;
                        }
                    }
                });
        }
        if (!done) {
            int have = klo + cin;
            if (have <= a.kth) {
                // too few values up to the probe: move the bracket up
                lo2 = probe;
                klo = have;
                probe = up2 < INFINITY ? up2 : 4.0 * probe;
            } else if (cin <= KNN_CAP) {
                int rk = a.kth - klo;
                for (int c = 0; c < cin; ++c) {
                    double v = lst[c * 64 + lane];
                    int less = 0, leq = 0;
                    for (int d = 0; d < cin; ++d) {
                        double w = lst[d * 64 + lane];
                        less += w < v ? 1 : 0;
                        leq += w <= v ? 1 : 0;
                    }
                    if (less <= rk && rk < leq) res = v;
                }
                done = true;
            } else {
                // too many to keep: bisect the bracket in the ordering of
                // the IEEE bit patterns (monotonic for values >= 0), so ties
                // and tiny values resolve within 64 halvings
                up2 = probe;
                unsigned long long lb = (unsigned long long)__double_as_longlong(lo2 < 0.0 ? 0.0 : lo2);
                unsigned long long pb = (unsigned long long)__double_as_longlong(probe);
                double mid = __longlong_as_double((long long)(lb + (pb - lb) / 2));
                if (mid > lo2 && mid < probe) {
                    probe = mid;
                } else {
                    res = probe;  // every value left in the bracket is the same
                    done = true;
                }
            }
        }
    }
    if (ok) {
        if (!done) atomicAdd((unsigned long long *)&a.counters[CNT_H_CAPPED], 1ull);
        s.h[i] = done ? 0.5 * sqrt(res) : NAN;
    }
}

// Neighbour lists under r_ij^2 <= (h_i + h_j)^2 through the bucket culling.
// counts == nullptr: fill mode, writing the neighbours' caller ids (s.id) at
// offsets[id_i - r0] for targets with r0 <= id_i < r1.
__global__ __launch_bounds__(64) void k_nbr(Tables t, Soa s, int64_t r0, int64_t r1,
                                            int32_t *__restrict__ counts,
                                            const int64_t *__restrict__ offsets,
                                            int32_t *__restrict__ idx) {
    __shared__ double sx[64], sy[64], sz[64], sh[64];
    __shared__ int32_t sid[64];
    int b = t.boff + xcd_block(blockIdx.x, gridDim.x);
    int lane = threadIdx.x;
    int cnt = t.b_cnt[b];
    int i = t.b_beg[b] + lane;
    bool ok = lane < cnt;
    double xi = ok ? s.x[i] : 0.0, yi = ok ? s.y[i] : 0.0, zi = ok ? s.z[i] : 0.0;
    double hi = ok ? s.h[i] : 0.0;
    int32_t myid = ok ? s.id[i] : -1;
    bool want = ok && myid >= r0 && myid < r1;
    int64_t o = (want && !counts) ? offsets[myid - r0] : 0;
    double hI = t.bhmax[b];
    const double ib[6] = {t.bxl[b], t.bxh[b], t.byl[b], t.byh[b], t.bzl[b], t.bzh[b]};
    int c = 0;
    for_candidates(
        t, t.b_run[b], ib, lane, [&](int S) { return hI + t.shmax[S]; },
        [&](int J) { return hI + t.bhmax[J]; },
        [&](int J) {
            int jc = t.b_cnt[J], j0 = t.b_beg[J];
            __syncthreads();
            if (lane < jc) {
                sx[lane] = s.x[j0 + lane];
                sy[lane] = s.y[j0 + lane];
                sz[lane] = s.z[j0 + lane];
                sh[lane] = s.h[j0 + lane];
                sid[lane] = s.id[j0 + lane];
            }
            __syncthreads();
            for (int q = 0; q < jc; ++q) {
                double r2 = sphx_r2(xi - sx[q], yi - sy[q], zi - sz[q]);
                double tt = hi + sh[q];
                if (want && r2 <= sphx_sq(tt)) {
                    if (!counts) idx[o + c] = sid[q];
                    ++c;
                }
            }
        });
    if (want && counts) counts[myid - r0] = c;
}

}  // namespace

int sphx_knn_pass(Engine &E, int kth, hipStream_t st) {
    if (E.n == 0) return SPHX_OK;
    if (E.nb_launch() == 0) return SPHX_OK;
    KnnArgs a{kth, 200, E.counters};
    k_knn<<<E.nb_launch(), 64, 0, st>>>(E.tview(), E.s, a);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int sphx_neighbor_pass(Engine &E, int64_t r0, int64_t r1, int32_t *counts, const int64_t *offsets,
                       int32_t *idx, hipStream_t st) {
    if (E.n == 0) return SPHX_OK;
    k_nbr<<<E.t.nb, 64, 0, st>>>(E.t, E.s, r0, r1, counts, offsets, idx);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

// Small runs walk half a bucket per group: with one wave per bucket most
// SIMDs would hold a single wave, and the walks wait on their loads.  Up to
// 1,536 buckets for the h iteration and 3,072 for density (measured: at
// 65,752 particles smoothing 0.21 -> 0.15 ms and density 0.33 -> 0.22 ms, at
// 131,072 density 0.30 -> 0.26 ms; smoothing at 131,072 and both at 262,144
// were slower with halves).
constexpr unsigned SMOOTH_HALF_BUCKETS = 1536, DENS_HALF_BUCKETS = 3072;

int Engine::walk_halves() const { return nb_launch() <= DENS_HALF_BUCKETS ? 2 : 1; }

int Engine::smoothing(hipStream_t st, bool predictor) {
    if (n == 0 || nb_launch() == 0) return SPHX_OK;
    const int halves = nb_launch() <= SMOOTH_HALF_BUCKETS ? 2 : 1;
    SmoothArgs a{p.nngb, p.ndown, p.nup, p.max_iter, (predictor && p.h_predictor) ? 1 : 0,
                 d_run, h_iters, counters, halves, sphx_h_step(p.h_mode), p.h_min};
    if (p.h_mode == 5) k_smoothing<5><<<nb_launch() * halves, 64, 0, st>>>(tview(), s, a);
    else if (p.h_mode == 4) k_smoothing<4><<<nb_launch() * halves, 64, 0, st>>>(tview(), s, a);
    else if (p.h_mode >= 2) k_smoothing<2><<<nb_launch() * halves, 64, 0, st>>>(tview(), s, a);
    else k_smoothing<0><<<nb_launch() * halves, 64, 0, st>>>(tview(), s, a);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

bool Engine::lists_ready() const {
    return use_lists && nbr && prec && fbal && nbr_ovf && nbr_bcap >= t.nb;
}

int Engine::density_eos(hipStream_t st, bool lists) {
    if (n == 0 || nb_launch() == 0) return SPHX_OK;
    const int halves = walk_halves();
    DensArgs a{halves, p.gamma, tmpi, counters, nullptr, nullptr, nullptr};
    if (lists) {
        if (!lists_ready()) {
            sphx_set_error("neighbour lists requested but not allocated");
            return SPHX_ERR_STATE;
        }
        a.lst = nbr;
        a.rec = prec;
        a.ovf = nbr_ovf;
    }
    k_density<<<nb_launch() * halves, 64, 0, st>>>(tview(), s, a);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

// Launch shape of the overflow walks: enough waves to spread the flag scan,
// never more than one per 64 buckets.
static unsigned ovf_grid(unsigned nb) { return std::min(256u, (nb + 63) / 64); }

int Engine::divcurl(hipStream_t st, bool lists) {
    if (n == 0 || nb_launch() == 0) return SPHX_OK;
    if (lists) {
        ListArgs L{nbr, prec, nbr_ovf, fbal};
        k_divcurl_list<<<nb_launch(), 64, 0, st>>>(tview(), s, L);
        k_divcurl_ovf<<<ovf_grid(nb_launch()), 64, 0, st>>>(tview(), s, nbr_ovf,
                                                           (int)nb_launch());
        if (p.balsara) {
            int64_t p0 = rrun >= 0 ? rp0 : 0, p1 = rrun >= 0 ? rp1 : n;
            if (p1 > p0)
                k_fbal<<<(unsigned)((p1 - p0 + 255) / 256), 256, 0, st>>>(s, p0, p1, fbal);
        }
    } else {
        k_divcurl<<<nb_launch(), 64, 0, st>>>(tview(), s);
    }
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int Engine::divcurl_hydro_lists(hipStream_t st) {
    if (n == 0 || nb_launch() == 0) return SPHX_OK;
    if (p.balsara || !lists_ready()) {
        sphx_set_error("the fused list pass needs the lists and no Balsara switch");
        return SPHX_ERR_STATE;
    }
    HydroArgs a{(float)p.alpha, (float)p.beta, (float)(p.eta * p.eta), (float)p.av_w_eps,
                p.av_form, p.balsara};
    ListArgs L{nbr, prec, nbr_ovf, fbal};
    const int nb = (int)nb_launch();
    k_dchydro_list<<<nb, 64, 0, st>>>(tview(), s, a, L);
    k_divcurl_ovf<<<ovf_grid(nb), 64, 0, st>>>(tview(), s, nbr_ovf, nb);
    k_hydro_ovf<<<ovf_grid(nb), 64, 0, st>>>(tview(), s, a, nbr_ovf, nb, 1);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int Engine::hydro(hipStream_t st, bool lists) {
    if (n == 0 || nb_launch() == 0) return SPHX_OK;
    HydroArgs a{(float)p.alpha, (float)p.beta, (float)(p.eta * p.eta), (float)p.av_w_eps,
                p.av_form, p.balsara};
    if (lists) {
        ListArgs L{nbr, prec, nbr_ovf, fbal};
        k_hydro_list<<<nb_launch(), 64, 0, st>>>(tview(), s, a, L);
        k_hydro_ovf<<<ovf_grid(nb_launch()), 64, 0, st>>>(tview(), s, a, nbr_ovf,
                                                         (int)nb_launch(), 0);
    } else {
        k_hydro<<<nb_launch(), 64, 0, st>>>(tview(), s, a);
    }
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}
