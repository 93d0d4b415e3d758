// C ABI of libsphx: engine lifecycle, run tables, step orchestration, the
// stateless drop-in calls and the resident plan.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>

#include "sphx_internal.h"

int sphx_zero_counters(int64_t *c, hipStream_t st);
int sphx_combine(Engine &E, hipStream_t st);
int sphx_neighbor_pass(Engine &E, int64_t r0, int64_t r1, int32_t *counts, const int64_t *offsets,
                       int32_t *idx, hipStream_t st);
CoolDev sphx_engine_cool(const Engine &E);

static thread_local std::string g_err;
void sphx_set_error(const std::string &msg) { g_err = msg; }

extern "C" const char *sphx_last_error(void) { return g_err.c_str(); }
extern "C" const char *sphx_version(void) { return "sphx 0.1 gfx950"; }

extern "C" void sphx_default_params(sphx_params *p) {
    p->gamma = 5.0 / 3.0;
    p->alpha = 1.0;
    p->beta = 2.0;
    p->eta = 0.1;
    p->G = 1.0;
    p->nngb = 64;
    p->ndown = 59;
    p->nup = 69;
    p->max_iter = 100;
    p->av_form = 0;
    p->balsara = 0;
    p->eps_from_h = 1;
    p->h_predictor = 1;
    p->h_mode = 0;
    p->knn_kth = 64;
    p->av_w_eps = 1e-5;
    p->h_min = 0.0;
}

namespace {

constexpr int UT = 256;
inline unsigned gu(int64_t n) { return (unsigned)((n + UT - 1) / UT); }

__global__ void k_iota(int32_t *a, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * UT + threadIdx.x;
    if (i < n) a[i] = (int32_t)i;
}
__global__ void k_fill_i(int32_t *a, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * UT + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_fill_d(double *a, double v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * UT + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_copy_d(const double *__restrict__ a, double *__restrict__ b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * UT + threadIdx.x;
    if (i < n) b[i] = a[i];
}
__global__ void k_copy_i(const int32_t *__restrict__ a, int32_t *__restrict__ b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * UT + threadIdx.x;
    if (i < n) b[i] = a[i];
}
// dst[i] = src[id[i]] (bring a caller array into the engine's order)
__global__ void k_gather_by_id(const double *__restrict__ src, const int32_t *__restrict__ id,
                               double *__restrict__ dst, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * UT + threadIdx.x;
    if (i < n) dst[i] = src[id[i]];
}
// (n, 3) rows in caller order -> SoA in engine order: slot k takes row id[k].
__global__ void k_gather_aos_by_id(const double *__restrict__ aos, const int32_t *__restrict__ id,
                                   double *__restrict__ x, double *__restrict__ y,
                                   double *__restrict__ z, int64_t n) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int64_t q = 3 * (int64_t)id[k];
    x[k] = aos[q];
    y[k] = aos[q + 1];
    z[k] = aos[q + 2];
}
__global__ void k_gather_by_id_i(const int32_t *__restrict__ src, const int32_t *__restrict__ id,
                                 int32_t *__restrict__ dst, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * UT + threadIdx.x;
    if (i < n) dst[i] = src[id[i]];
}
// out[id[i] - r0] = val[i] for r0 <= id[i] < r1
__global__ void k_scatter_d(const double *__restrict__ val, const int32_t *__restrict__ id,
                            int64_t r0, int64_t r1, double *__restrict__ out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * UT + threadIdx.x;
    if (i >= n) return;
    int64_t o = id[i];
    if (o >= r0 && o < r1) out[o - r0] = val[i];
}
__global__ void k_scatter_i(const int32_t *__restrict__ val, const int32_t *__restrict__ id,
                            int64_t r0, int64_t r1, int32_t *__restrict__ out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * UT + threadIdx.x;
    if (i >= n) return;
    int64_t o = id[i];
    if (o >= r0 && o < r1) out[o - r0] = val[i];
}
__global__ void k_scatter_aos(const double *__restrict__ a, const double *__restrict__ b,
                              const double *__restrict__ c, const int32_t *__restrict__ id,
                              int64_t r0, int64_t r1, double *__restrict__ out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * UT + threadIdx.x;
    if (i >= n) return;
    int64_t o = id[i];
    if (o >= r0 && o < r1) {
        int64_t q = 3 * (o - r0);
        out[q] = a[i];
        out[q + 1] = b[i];
        out[q + 2] = c[i];
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Engine lifecycle
// ---------------------------------------------------------------------------
Engine::~Engine() {
    int cur = 0;
    hipGetDevice(&cur);
    hipSetDevice(device);
    if (bh) sphx_bh_host_free(bh);
    if (grav.sym) sphx_sym_host_free(grav.sym);
    for (void *q : owned) hipFree(q);
    for (hipEvent_t e : tev)
        if (e) hipEventDestroy(e);
    drop_graph();
    if (gev_in) hipEventDestroy(gev_in);
    if (gev_out) hipEventDestroy(gev_out);
    if (gstream) hipStreamDestroy(gstream);
    if (cool_T) hipFree(cool_T);
    if (cool_L) hipFree(cool_L);
    if (grid_buf) hipFree(grid_buf);
    if (chimes_buf) hipFree(chimes_buf);
    if (chimes_ion) hipFree(chimes_ion);
    hipSetDevice(cur);
}

int Engine::alloc(void **ptr, size_t bytes) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes ? bytes : 8);
    if (e != hipSuccess) {
        sphx_set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
        return SPHX_ERR_NOMEM;
    }
    owned.push_back(q);
    *ptr = q;
    return SPHX_OK;
}

int Engine::reserve(int64_t capacity) {
    if (capacity <= cap) return SPHX_OK;
    for (void *q : owned) hipFree(q);
    owned.clear();
    grav = GravHost{};
    tab_i32 = nullptr; tab_i32_cap = 0;
    tab_d = nullptr; tab_d_cap = 0;
    bbox = nullptr; bbox_cap = 0;
    d_run = nullptr;
    dtpart = nullptr; dtpart_cap = 0;
    nbr = nullptr; nbr_ovf = nullptr; nbr_bcap = 0; prec = nullptr; fbal = nullptr;
    cap = 0;
    n = 0;
    R = 0;
    run_n.clear();
    size_t D = sizeof(double) * (size_t)capacity, I = sizeof(int32_t) * (size_t)capacity;
    double **dptr[] = {&s.x, &s.y, &s.z, &s.vx, &s.vy, &s.vz, &s.u, &s.m, &s.h, &s.eps,
                       &s.ut_prev, &s.ax, &s.ay, &s.az, &s.gx, &s.gy, &s.gz, &s.rho, &s.P, &s.c,
                       &s.divv, &s.curlv, &s.du, &s.pot};
    for (double **q : dptr) SPHX_CHECK(alloc((void **)q, D));
    // tmpd also holds the 256 x 11 diagnostic block partials
    SPHX_CHECK(alloc((void **)&tmpd, sizeof(double) * (size_t)std::max<int64_t>(capacity, 4096)));
    int32_t **iptr[] = {&s.k, &s.id, &s.run, &s.bin, &bin_spare, &perm, &perm2, &tmpi, &h_iters};
    for (int32_t **q : iptr) SPHX_CHECK(alloc((void **)q, I));
    SPHX_CHECK(alloc((void **)&keys, sizeof(uint64_t) * capacity));
    SPHX_CHECK(alloc((void **)&keys2, sizeof(uint64_t) * capacity));
    hist_cap = 256 * ((capacity + SPHX_RS_TILE - 1) / SPHX_RS_TILE + 1);
    SPHX_CHECK(alloc((void **)&hist, sizeof(uint32_t) * hist_cap));
    SPHX_CHECK(alloc((void **)&counters, sizeof(int64_t) * CNT_SLOTS));
    SPHX_HIP(hipMemset(counters, 0, sizeof(int64_t) * CNT_SLOTS));
    if (want_lists && mode == SPHX_MIXED) {
        // every run has at most one partial bucket
        int64_t bc = capacity / SPHX_BUCKET + std::max(max_runs, 1) + 1;
        SPHX_CHECK(alloc((void **)&nbr, sizeof(int32_t) * (size_t)bc * SPHX_NBR_MAX * 64));
        SPHX_CHECK(alloc((void **)&nbr_ovf, sizeof(int32_t) * 2 * (size_t)bc));
        SPHX_CHECK(alloc((void **)&prec, sizeof(PairRec) * (size_t)capacity));
        SPHX_CHECK(alloc((void **)&fbal, sizeof(float) * (size_t)capacity));
        nbr_bcap = bc;
    }
    cap = capacity;
    return SPHX_OK;
}

int Engine::set_runs(const std::vector<int64_t> &sizes, hipStream_t st) {
    int64_t total = 0;
    for (int64_t v : sizes) total += v;
    if (total > cap) {
        sphx_set_error("run sizes exceed the engine capacity");
        return SPHX_ERR_ARG;
    }
    if (total > INT32_MAX) {
        sphx_set_error("more than 2^31-1 particles");
        return SPHX_ERR_ARG;
    }
    hb_beg.clear(); hb_cnt.clear(); hb_run.clear(); hs_bbeg.clear(); hs_bend.clear();
    hr_sbeg.clear(); hr_send.clear(); hr_pbeg.clear(); hr_pend.clear(); hr_lcut.clear();
    int64_t off = 0;
    for (size_t r = 0; r < sizes.size(); ++r) {
        int64_t nr = sizes[r];
        hr_pbeg.push_back((int32_t)off);
        hr_pend.push_back((int32_t)(off + nr));
        // A bucket is cut where consecutive particles lie in different cells
        // of this Morton level: cells of about 8 x 64 particles or more, so
        // the cut costs few extra segments but a bucket never straddles a big
        // jump of the curve.  floor(log8(n / 64)) - 1, at least 1.
        {
            int l = 0;
            while (((int64_t)64 << (3 * (l + 1))) <= nr) ++l;
            hr_lcut.push_back(std::max(1, l - 1 + lcut_bias));
        }
        int b0 = (int)hb_beg.size();
        for (int64_t a = 0; a < nr; a += SPHX_BUCKET) {
            hb_beg.push_back((int32_t)(off + a));
            hb_cnt.push_back((int32_t)std::min<int64_t>(SPHX_BUCKET, nr - a));
            hb_run.push_back((int32_t)r);
        }
        int b1 = (int)hb_beg.size();
        hr_sbeg.push_back((int32_t)hs_bbeg.size());
        for (int b = b0; b < b1; b += SPHX_SUPER) {
            hs_bbeg.push_back(b);
            hs_bend.push_back(std::min(b1, b + SPHX_SUPER));
        }
        hr_send.push_back((int32_t)hs_bbeg.size());
        off += nr;
    }
    int nb = (int)hb_beg.size(), ns = (int)hs_bbeg.size(), nr = (int)sizes.size();
    int64_t ni = 3 * (int64_t)nb + 2 * (int64_t)ns + 5 * (int64_t)nr;
    if (ni > tab_i32_cap) {
        SPHX_CHECK(alloc((void **)&tab_i32, sizeof(int32_t) * (ni + 1024)));
        tab_i32_cap = ni + 1024;
    }
    std::vector<int32_t> all;
    all.reserve(ni);
    for (auto *v : {&hb_beg, &hb_cnt, &hb_run, &hs_bbeg, &hs_bend, &hr_sbeg, &hr_send, &hr_pbeg,
                    &hr_pend, &hr_lcut})
        all.insert(all.end(), v->begin(), v->end());
    if (!all.empty()) {
        SPHX_HIP(hipMemcpyAsync(tab_i32, all.data(), sizeof(int32_t) * all.size(),
                                hipMemcpyHostToDevice, st));
        SPHX_HIP(hipStreamSynchronize(st));  // `all` is a host temporary
    }
    int64_t nd = 8 * (int64_t)nb + 7 * (int64_t)ns;
    if (nd > tab_d_cap) {
        SPHX_CHECK(alloc((void **)&tab_d, sizeof(double) * (nd + 1024)));
        tab_d_cap = nd + 1024;
    }
    int64_t nbx = 6 * (int64_t)ns + 6 * (int64_t)nr;
    if (nbx > bbox_cap) {
        SPHX_CHECK(alloc((void **)&bbox, sizeof(double) * (nbx + 1024)));
        bbox_cap = nbx + 1024;
    }
    if ((int)run_params.size() < nr) run_params.resize(nr, RunParams{1e-4, 0, p.eps_from_h});
    if (5 * (int64_t)nb > dtpart_cap) {
        SPHX_CHECK(alloc((void **)&dtpart, sizeof(double) * (5 * (int64_t)nb + 1024)));
        dtpart_cap = 5 * (int64_t)nb + 1024;
    }
    if (R < nr || !d_run) {
        SPHX_CHECK(alloc((void **)&d_run, sizeof(RunParams) * std::max(nr, 1)));
    }
    drop_graph();
    int32_t *q = tab_i32;
    t.nb = nb; t.ns = ns; t.R = nr;
    t.boff = 0;
    t.b_beg = q; q += nb;
    t.b_cnt = q; q += nb;
    t.b_run = q; q += nb;
    t.s_bbeg = q; q += ns;
    t.s_bend = q; q += ns;
    t.r_sbeg = q; q += nr;
    t.r_send = q; q += nr;
    t.r_pbeg = q; q += nr;
    t.r_pend = q; q += nr;
    t.r_lcut = q; q += nr;
    double *d = tab_d;
    double **bx[] = {&t.bxl, &t.bxh, &t.byl, &t.byh, &t.bzl, &t.bzh, &t.bhmax, &t.bemax};
    for (double **a : bx) { *a = d; d += nb; }
    double **sx[] = {&t.sxl, &t.sxh, &t.syl, &t.syh, &t.szl, &t.szh, &t.shmax};
    for (double **a : sx) { *a = d; d += ns; }
    run_n = sizes;
    R = nr;
    n = total;
    return SPHX_OK;
}

int Engine::upload_run_params(hipStream_t st) {
    if (R == 0) return SPHX_OK;
    SPHX_HIP(hipMemcpyAsync(d_run, run_params.data(), sizeof(RunParams) * R, hipMemcpyHostToDevice, st));
    return SPHX_OK;
}

int Engine::pull_run_params(hipStream_t st) {
    if (R == 0 || !d_run) return SPHX_OK;
    if ((int)run_params.size() < R) run_params.resize(R);
    SPHX_HIP(hipMemcpyAsync(run_params.data(), d_run, sizeof(RunParams) * R, hipMemcpyDeviceToHost, st));
    SPHX_HIP(hipStreamSynchronize(st));
    return SPHX_OK;
}

int Engine::tmark(hipStream_t st, int phase, int end) {
    if (!timing) return SPHX_OK;
    hipEvent_t &e = tev[2 * phase + end];
    if (!e) SPHX_HIP(hipEventCreate(&e));
    SPHX_HIP(hipEventRecord(e, st));
    if (end) tmarked[phase] = true;
    return SPHX_OK;
}

int Engine::tcollect() {
    if (!timing) return SPHX_OK;
    for (int ph = 0; ph < PH_COUNT; ++ph) {
        hipEvent_t a = tev[2 * ph], b = tev[2 * ph + 1];
        if (!a || !b || !tmarked[ph]) continue;
        tmarked[ph] = false;
        SPHX_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, a, b) == hipSuccess) {
            phase_ms[ph] += ms;
            phase_n[ph] += 1;
        }
    }
    return SPHX_OK;
}

int Engine::forces(hipStream_t st, bool predictor, bool with_pot) {
    SPHX_CHECK(tmark(st, PH_SMOOTH, 0));
    SPHX_CHECK(build_boxes(st, false));
    if (p.h_mode == 1) {
        // kNN h (do_smoothingX): a count iteration brackets the radius of rank
        // knn_kth, then the selection pass makes it exact
        sphx_params saved = p;
        p.nngb = saved.knn_kth + 16;
        p.ndown = saved.knn_kth + 1;
        p.nup = saved.knn_kth + 56;
        int rc = smoothing(st, false);
        p = saved;
        SPHX_CHECK(rc);
        SPHX_CHECK(sphx_knn_pass(*this, p.knn_kth, st));
    } else {
        SPHX_CHECK(smoothing(st, predictor));
    }
    SPHX_CHECK(tmark(st, PH_SMOOTH, 1));
    // the density pass writes neighbour lists that the div/curl and hydro
    // passes of this step iterate (positions and h do not change in between)
    const bool lists = lists_ready();
    SPHX_CHECK(tmark(st, PH_DENS, 0));
    SPHX_CHECK(build_boxes(st, true));
    SPHX_CHECK(density_eos(st, lists));
    SPHX_CHECK(tmark(st, PH_DENS, 1));
    SPHX_CHECK(tmark(st, PH_GRAV, 0));
    if (bh_theta > 0.0) {
        if (!bh) bh = sphx_bh_host_new();
        SPHX_CHECK(sphx_bh_run(*this, *bh, bh_theta, with_pot, st, nullptr));
    } else {
        SPHX_CHECK(sphx_gravity_run(*this, gcfg, with_pot, st));
    }
    SPHX_CHECK(tmark(st, PH_GRAV, 1));
    // without the Balsara switch hydro reads nothing of div/curl's output, and
    // one pass over the lists gives both and the total acceleration (timed as
    // the div/curl phase)
    const bool fused = lists && !p.balsara;
    SPHX_CHECK(tmark(st, PH_DIVCURL, 0));
    SPHX_CHECK(fused ? divcurl_hydro_lists(st) : divcurl(st, lists));
    SPHX_CHECK(tmark(st, PH_DIVCURL, 1));
    SPHX_CHECK(tmark(st, PH_HYDRO, 0));
    if (!fused) SPHX_CHECK(hydro(st, lists));
    SPHX_CHECK(tmark(st, PH_HYDRO, 1));
    // the fused pass adds gravity to the accelerations as it stores them
    if (!fused) SPHX_CHECK(sphx_combine(*this, st));
    return SPHX_OK;
}

int Engine::plain_forces(hipStream_t st, bool predictor, bool with_pot) {
    if (p.h_mode == 1) {
        sphx_set_error("kNN smoothing lengths (h_mode 1) are a mixed-precision option");
        return SPHX_ERR_ARG;
    }
    bool pred = predictor && p.h_predictor;
    // the step's dt is read on the device (run 0's parameters), where the
    // adaptive time step leaves it
    const double *dt_p = d_run ? &d_run[0].dt : nullptr;
    double dt = run_params.empty() ? 0.0 : run_params[0].dt;
    SPHX_CHECK(tmark(st, PH_SMOOTH, 0));
    SPHX_CHECK(plain_smoothing(s.x, s.y, s.z, s.h, pred ? s.k : nullptr, pred ? s.divv : nullptr,
                               dt_p, dt, nullptr, p, n, 0, n, s.h, s.k, h_iters, counters, st));
    SPHX_CHECK(tmark(st, PH_SMOOTH, 1));
    SPHX_CHECK(tmark(st, PH_DENS, 0));
    SPHX_CHECK(plain_density(s.x, s.y, s.z, s.m, s.h, n, 0, n, s.rho, tmpi, st));
    SPHX_CHECK(plain_eos(s.rho, s.u, p.gamma, n, s.P, s.c, st));
    SPHX_CHECK(tmark(st, PH_DENS, 1));
    const double *eps = p.eps_from_h ? s.h : s.eps;
    SPHX_CHECK(tmark(st, PH_GRAV, 0));
    SPHX_CHECK(plain_gravity(s.x, s.y, s.z, s.m, eps, p.G, n, 0, n, s.gx, s.gy, s.gz,
                             with_pot ? s.pot : nullptr, st));
    SPHX_CHECK(tmark(st, PH_GRAV, 1));
    SPHX_CHECK(tmark(st, PH_DIVCURL, 0));
    SPHX_CHECK(plain_divcurl(s.x, s.y, s.z, s.vx, s.vy, s.vz, s.rho, s.m, s.h, n, 0, n, s.divv,
                             s.curlv, st));
    SPHX_CHECK(tmark(st, PH_DIVCURL, 1));
    SPHX_CHECK(tmark(st, PH_HYDRO, 0));
    SPHX_CHECK(plain_acc_sph(s, p, n, 0, n, s.ax, s.ay, s.az, st));
    SPHX_CHECK(plain_du(s, p, n, 0, n, s.du, st));
    SPHX_CHECK(tmark(st, PH_HYDRO, 1));
    SPHX_CHECK(sphx_combine(*this, st));
    return SPHX_OK;
}

int Engine::step_body(hipStream_t st) {
    SPHX_CHECK(sphx_zero_counters(counters, st));
    SPHX_CHECK(tmark(st, PH_SORT, 0));
    SPHX_CHECK(kick_drift(st));
    if (mode == SPHX_MIXED) SPHX_CHECK(sort_and_permute(st, true));
    SPHX_CHECK(tmark(st, PH_SORT, 1));
    if (mode == SPHX_MIXED) SPHX_CHECK(forces(st, true, false));
    else SPHX_CHECK(plain_forces(st, true, false));
    SPHX_CHECK(tmark(st, PH_FINISH, 0));
    SPHX_CHECK(finish_step(st));
    SPHX_CHECK(global_dt(st));
    SPHX_CHECK(tmark(st, PH_FINISH, 1));
    return SPHX_OK;
}

// One substep of the block time steps (dt_max / 2^B): opening kicks at the
// first substep, drift of every particle, sort, forces (gravity only for the
// blocks holding a particle whose step ends here; the SPH passes for all, as
// the neighbours of active particles need their current h, rho, P, c),
// closing and opening kicks of the active particles, t += substep.
int Engine::block_substep(hipStream_t st) {
    SPHX_CHECK(sphx_zero_counters(counters, st));
    SPHX_CHECK(tmark(st, PH_SORT, 0));
    if (!blk_started) {
        SPHX_CHECK(block_open(st));
        blk_started = true;
    }
    SPHX_CHECK(block_drift(std::ldexp(blk_dtmax, -blk_B), st));
    SPHX_CHECK(sort_and_permute(st, true));
    SPHX_CHECK(tmark(st, PH_SORT, 1));
    int64_t t1 = blk_tick + 1;
    blk_t1 = t1;
    int rc = forces(st, true, false);
    blk_t1 = 0;
    SPHX_CHECK(rc);
    SPHX_CHECK(tmark(st, PH_FINISH, 0));
    SPHX_CHECK(block_close(t1, st));
    SPHX_CHECK(global_dt(st));  // t += the substep (run dt)
    SPHX_CHECK(tmark(st, PH_FINISH, 1));
    blk_tick = t1 & ((1ll << blk_B) - 1);
    return SPHX_OK;
}

int Engine::step(hipStream_t st, int nsteps) {
    for (int k = 0; k < nsteps; ++k) {
        SPHX_CHECK(blk_B > 0 ? block_substep(st) : step_body(st));
        SPHX_CHECK(tcollect());
        ++steps;
    }
    return SPHX_OK;
}

void Engine::drop_graph() {
    for (int p = 0; p < 2; ++p) {
        if (graph_exec[p]) hipGraphExecDestroy(graph_exec[p]);
        if (graph[p]) hipGraphDestroy(graph[p]);
        graph_exec[p] = nullptr;
        graph[p] = nullptr;
        graph_valid[p] = false;
    }
}

// One step through the captured graph of the current pointer parity.  The
// first call per parity captures the step (the host-side pointer swap of the
// sort happens during capture); later calls replay it and mirror the swap.
int Engine::step_graph(hipStream_t st) {
    if (mode != SPHX_MIXED) return step(st, 1);
    int p = gpar;
    if (!graph_valid[p]) {
        if (graph_exec[p]) hipGraphExecDestroy(graph_exec[p]);
        if (graph[p]) hipGraphDestroy(graph[p]);
        graph_exec[p] = nullptr;
        graph[p] = nullptr;
        SPHX_CHECK(sphx_gravity_prepare(*this, gcfg, st));
        SPHX_HIP(hipStreamSynchronize(st));
        bool t_on = timing;
        timing = false;
        SPHX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = step_body(st);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g);
        timing = t_on;
        if (rc < 0 || e != hipSuccess) {
            if (g) hipGraphDestroy(g);
            if (rc >= 0) {
                sphx_set_error(std::string("stream capture: ") + hipGetErrorString(e));
                rc = SPHX_ERR_HIP;
            }
            return rc;
        }
        hipError_t ei = hipGraphInstantiate(&graph_exec[p], g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            hipGraphDestroy(g);
            sphx_set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
            return SPHX_ERR_HIP;
        }
        graph[p] = g;
        graph_valid[p] = true;
        // capture ran the host side of the step (gpar is now p ^ 1); execute it
        SPHX_HIP(hipGraphLaunch(graph_exec[p], st));
    } else {
        SPHX_HIP(hipGraphLaunch(graph_exec[p], st));
        swap_sorted();
    }
    ++steps;
    return SPHX_OK;
}

int Engine::step_graphs(hipStream_t caller, int nsteps) {
    if (!gstream) SPHX_HIP(hipStreamCreateWithFlags(&gstream, hipStreamNonBlocking));
    if (!gev_in) SPHX_HIP(hipEventCreateWithFlags(&gev_in, hipEventDisableTiming));
    if (!gev_out) SPHX_HIP(hipEventCreateWithFlags(&gev_out, hipEventDisableTiming));
    SPHX_HIP(hipEventRecord(gev_in, caller));
    SPHX_HIP(hipStreamWaitEvent(gstream, gev_in, 0));
    for (int k = 0; k < nsteps; ++k) SPHX_CHECK(step_graph(gstream));
    SPHX_HIP(hipEventRecord(gev_out, gstream));
    SPHX_HIP(hipStreamWaitEvent(caller, gev_out, 0));
    return SPHX_OK;
}

int Engine::restrict_to(int run) {
    if (run < 0) {
        rrun = -1;
        return SPHX_OK;
    }
    if (run >= R) {
        sphx_set_error("run index out of range");
        return SPHX_ERR_ARG;
    }
    rrun = run;
    rp0 = hr_pbeg[run];
    rp1 = hr_pend[run];
    if (hr_send[run] > hr_sbeg[run]) {
        rb0 = hs_bbeg[hr_sbeg[run]];
        rb1 = hs_bend[hr_send[run] - 1];
    } else {
        rb0 = rb1 = 0;
    }
    return SPHX_OK;
}

int Engine::read_counters(sphx_counters *out) {
    int64_t c[CNT_SLOTS];
    SPHX_HIP(hipMemcpy(c, counters, sizeof(c), hipMemcpyDeviceToHost));
    out->n = n;
    out->steps = steps;
    out->h_iters_max = c[CNT_H_ITERS_MAX];
    out->h_iters_sum = c[CNT_H_ITERS_SUM];
    out->h_capped = c[CNT_H_CAPPED];
    out->cool_capped = c[CNT_COOL_CAPPED];
    out->cand_buckets = c[CNT_CAND];
    out->buckets = t.nb;
    out->bh_overflow = c[CNT_BH_OVERFLOW];
    out->nonfinite = c[CNT_NONFINITE];
    out->bin_floor = c[CNT_BIN_FLOOR];
    out->cool_iters = c[CNT_COOL_ITERS];
    out->cool_bisected = c[CNT_COOL_BISECT];
    out->nbr_overflow = c[CNT_NBR_OVF];
    return SPHX_OK;
}

int Engine::diagnostics(hipStream_t st, bool with_pe, sphx_diag *out) {
    if (with_pe) {
        if (mode == SPHX_MIXED) {
            SPHX_CHECK(build_boxes(st, true));
            SPHX_CHECK(sphx_gravity_run(*this, gcfg, true, st));
        } else {
            const double *eps = p.eps_from_h ? s.h : s.eps;
            SPHX_CHECK(plain_gravity(s.x, s.y, s.z, s.m, eps, p.G, n, 0, n, s.gx, s.gy, s.gz,
                                     s.pot, st));
        }
    }
    return sphx_diag_reduce(s.x, s.y, s.z, s.vx, s.vy, s.vz, s.m, s.u, with_pe ? s.pot : nullptr, n,
                            tmpd, out, st);
}

// ---------------------------------------------------------------------------
// Argument checks
// ---------------------------------------------------------------------------
static int check_rows(int64_t n, int64_t r0, int64_t r1) {
    if (n < 0 || r0 < 0 || r1 < r0 || r1 > n) {
        sphx_set_error("bad particle count or row range");
        return SPHX_ERR_ARG;
    }
    return SPHX_OK;
}

static int check_mode(int mode) {
    if (mode != SPHX_MIXED && mode != SPHX_FP64) {
        sphx_set_error("mode must be SPHX_MIXED or SPHX_FP64");
        return SPHX_ERR_ARG;
    }
    return SPHX_OK;
}

static int check_params(const sphx_params *p) {
    if (!p || p->nngb <= 0 || p->ndown > p->nngb || p->nup < p->nngb || p->max_iter <= 0 ||
        !(p->gamma > 1.0) || (p->av_form != 0 && p->av_form != 1) ||
        p->h_mode < 0 || p->h_mode > 5 || p->knn_kth < 0 || p->knn_kth > 4096 ||
        !(p->av_w_eps >= 0.0) || !(p->h_min >= 0.0 && p->h_min < INFINITY)) {
        sphx_set_error("invalid sphx_params");
        return SPHX_ERR_ARG;
    }
    return SPHX_OK;
}

struct sphx_ctx {
    Engine E;
    BhHost *bh = nullptr;
    ~sphx_ctx() {
        if (bh) sphx_bh_host_free(bh);
    }
};

struct DeviceGuard {
    int prev = 0;
    explicit DeviceGuard(int dev) {
        hipGetDevice(&prev);
        hipSetDevice(dev);
    }
    ~DeviceGuard() { hipSetDevice(prev); }
};

// Load positions (+ optional SoA extras) into the context engine as a single
// run and, in mixed mode, put them in Morton order.  Afterwards s.id maps the
// engine's slots to the caller's indices.
static int ctx_load(sphx_ctx *ctx, int64_t n, const double *pos, const double *m, const double *h,
                    const double *v, bool sort, hipStream_t st) {
    Engine &E = ctx->E;
    SPHX_CHECK(E.reserve(std::max<int64_t>(n, 64)));
    SPHX_CHECK(E.set_runs({n}, st));
    if (n == 0) return SPHX_OK;
    SPHX_CHECK(sphx_aos_to_soa(pos, n, E.s.x, E.s.y, E.s.z, st));
    if (v) SPHX_CHECK(sphx_aos_to_soa(v, n, E.s.vx, E.s.vy, E.s.vz, st));
    if (m) k_copy_d<<<gu(n), UT, 0, st>>>(m, E.s.m, n);
    if (h) k_copy_d<<<gu(n), UT, 0, st>>>(h, E.s.h, n);
    k_iota<<<gu(n), UT, 0, st>>>(E.s.id, n);
    k_fill_i<<<gu(n), UT, 0, st>>>(E.s.run, 0, n);
    SPHX_HIP(hipGetLastError());
    if (sort) SPHX_CHECK(E.sort_and_permute(st, true));
    return SPHX_OK;
}

static int gather_in(Engine &E, const double *src, double *dst, hipStream_t st) {
    if (!src || E.n == 0) return SPHX_OK;
    k_gather_by_id<<<gu(E.n), UT, 0, st>>>(src, E.s.id, dst, E.n);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

static int scatter_out(Engine &E, const double *val, int64_t r0, int64_t r1, double *out,
                       hipStream_t st) {
    if (!out || E.n == 0) return SPHX_OK;
    k_scatter_d<<<gu(E.n), UT, 0, st>>>(val, E.s.id, r0, r1, out, E.n);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

static int scatter_out_i(Engine &E, const int32_t *val, int64_t r0, int64_t r1, int32_t *out,
                         hipStream_t st) {
    if (!out || E.n == 0) return SPHX_OK;
    k_scatter_i<<<gu(E.n), UT, 0, st>>>(val, E.s.id, r0, r1, out, E.n);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

static int scatter_out_aos(Engine &E, const double *a, const double *b, const double *c, int64_t r0,
                           int64_t r1, double *out, hipStream_t st) {
    if (!out || E.n == 0) return SPHX_OK;
    k_scatter_aos<<<gu(E.n), UT, 0, st>>>(a, b, c, E.s.id, r0, r1, out, E.n);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

// ---------------------------------------------------------------------------
// Stateless drop-ins
// ---------------------------------------------------------------------------
extern "C" int sphx_ctx_create(int device, sphx_ctx **out) {
    if (!out) {
        sphx_set_error("null output pointer");
        return SPHX_ERR_ARG;
    }
    int ndev = 0;
    SPHX_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) {
        sphx_set_error("no such HIP device");
        return SPHX_ERR_ARG;
    }
    sphx_ctx *c = new sphx_ctx();
    c->E.device = device;
    sphx_default_params(&c->E.p);
    *out = c;
    return SPHX_OK;
}

extern "C" int sphx_ctx_destroy(sphx_ctx *ctx) {
    if (!ctx) return SPHX_OK;
    delete ctx;
    return SPHX_OK;
}

#define CTX_PROLOGUE(ctx)                          \
    if (!ctx) {                                    \
        sphx_set_error("null context");            \
        return SPHX_ERR_ARG;                       \
    }                                              \
    DeviceGuard guard_(ctx->E.device);             \
    hipStream_t st = (hipStream_t)stream;

extern "C" int sphx_getDensity(sphx_ctx *ctx, int mode, int64_t n, const double *pos,
                               const double *m, const double *h, int64_t row_beg, int64_t row_end,
                               double *rho, int32_t *nngb, void *stream) {
    CTX_PROLOGUE(ctx);
    SPHX_CHECK(check_mode(mode));
    SPHX_CHECK(check_rows(n, row_beg, row_end));
    if (n == 0 || row_end == row_beg) return SPHX_OK;
    if (!pos || !m || !h || !rho) { sphx_set_error("null array"); return SPHX_ERR_ARG; }
    Engine &E = ctx->E;
    if (mode == SPHX_FP64) {
        SPHX_CHECK(ctx_load(ctx, n, pos, m, h, nullptr, false, st));
        return plain_density(E.s.x, E.s.y, E.s.z, E.s.m, E.s.h, n, row_beg, row_end, rho, nngb, st);
    }
    SPHX_CHECK(ctx_load(ctx, n, pos, m, h, nullptr, true, st));
    SPHX_CHECK(E.build_boxes(st, true));
    k_fill_d<<<gu(n), UT, 0, st>>>(E.s.u, 0.0, n);
    SPHX_CHECK(E.density_eos(st));
    SPHX_CHECK(scatter_out(E, E.s.rho, row_beg, row_end, rho, st));
    SPHX_CHECK(scatter_out_i(E, E.tmpi, row_beg, row_end, nngb, st));
    return SPHX_OK;
}

extern "C" int sphx_eos(sphx_ctx *ctx, int64_t n, const double *rho, const double *u, double gamma,
                        double *P, double *c, void *stream) {
    CTX_PROLOGUE(ctx);
    if (n < 0 || (n > 0 && (!rho || !u || !P || !c))) { sphx_set_error("bad arguments"); return SPHX_ERR_ARG; }
    return plain_eos(rho, u, gamma, n, P, c, st);
}

extern "C" int sphx_smoothing(sphx_ctx *ctx, int mode, int64_t n, const double *pos,
                              const double *h_prev, const int32_t *k_prev, const double *divV,
                              double dt, const sphx_params *p, int64_t row_beg, int64_t row_end,
                              double *h_out, int32_t *k_out, int32_t *iters_out, void *stream) {
    CTX_PROLOGUE(ctx);
    SPHX_CHECK(check_mode(mode));
    SPHX_CHECK(check_rows(n, row_beg, row_end));
    SPHX_CHECK(check_params(p));
    if (n == 0 || row_end == row_beg) return SPHX_OK;
    if (!pos || !h_prev || !h_out || !k_out) { sphx_set_error("null array"); return SPHX_ERR_ARG; }
    Engine &E = ctx->E;
    E.p = *p;
    bool pred = k_prev && divV;
    SPHX_CHECK(E.reserve(std::max<int64_t>(n, 64)));
    SPHX_CHECK(sphx_zero_counters(E.counters, st));
    if (mode == SPHX_FP64) {
        SPHX_CHECK(ctx_load(ctx, n, pos, nullptr, nullptr, nullptr, false, st));
        SPHX_CHECK(plain_smoothing(E.s.x, E.s.y, E.s.z, h_prev, pred ? k_prev : nullptr,
                                   pred ? divV : nullptr, nullptr, dt, nullptr, *p, n, row_beg,
                                   row_end, h_out, k_out, iters_out, E.counters, st));
    } else {
        SPHX_CHECK(ctx_load(ctx, n, pos, nullptr, h_prev, nullptr, true, st));
        if (pred) {
            SPHX_CHECK(gather_in(E, divV, E.s.divv, st));
            k_gather_by_id_i<<<gu(n), UT, 0, st>>>(k_prev, E.s.id, E.s.k, n);
            SPHX_HIP(hipGetLastError());
        }
        E.run_params.assign(1, RunParams{dt, 0, p->eps_from_h});
        SPHX_CHECK(E.upload_run_params(st));
        SPHX_CHECK(E.build_boxes(st, false));
        SPHX_CHECK(E.smoothing(st, pred));
        SPHX_CHECK(scatter_out(E, E.s.h, row_beg, row_end, h_out, st));
        SPHX_CHECK(scatter_out_i(E, E.s.k, row_beg, row_end, k_out, st));
        SPHX_CHECK(scatter_out_i(E, E.h_iters, row_beg, row_end, iters_out, st));
    }
    SPHX_HIP(hipStreamSynchronize(st));
    sphx_counters c{};
    SPHX_CHECK(E.read_counters(&c));
    return c.h_capped > 0 ? SPHX_CAPPED : SPHX_OK;
}

extern "C" int sphx_knn_h(sphx_ctx *ctx, int64_t n, const double *pos, const double *h_guess,
                          int32_t kth, int64_t row_beg, int64_t row_end, double *h_out,
                          void *stream) {
    CTX_PROLOGUE(ctx);
    SPHX_CHECK(check_rows(n, row_beg, row_end));
    if (n == 0 || row_end == row_beg) return SPHX_OK;
    if (!pos || !h_guess || !h_out) { sphx_set_error("null array"); return SPHX_ERR_ARG; }
    if (kth < 0 || kth >= n || kth > 4096) {
        sphx_set_error("kth must be in [0, min(n, 4097))");
        return SPHX_ERR_ARG;
    }
    Engine &E = ctx->E;
    sphx_params p;
    sphx_default_params(&p);
    // count band [kth + 1, kth + 56] around a target of kth + 16: the search
    // radius then holds at most 120 values, which the selection keeps in LDS
    p.nngb = kth + 16;
    p.ndown = kth + 1;
    p.nup = kth + 56;
    p.h_predictor = 0;
    E.p = p;
    SPHX_CHECK(E.reserve(std::max<int64_t>(n, 64)));
    SPHX_CHECK(sphx_zero_counters(E.counters, st));
    SPHX_CHECK(ctx_load(ctx, n, pos, nullptr, h_guess, nullptr, true, st));
    E.run_params.assign(1, RunParams{0.0, 0, p.eps_from_h});
    SPHX_CHECK(E.upload_run_params(st));
    SPHX_CHECK(E.build_boxes(st, false));
    SPHX_CHECK(E.smoothing(st, false));
    SPHX_CHECK(sphx_zero_counters(E.counters, st));
    SPHX_CHECK(sphx_knn_pass(E, kth, st));
    SPHX_CHECK(scatter_out(E, E.s.h, row_beg, row_end, h_out, st));
    SPHX_HIP(hipStreamSynchronize(st));
    sphx_counters c{};
    SPHX_CHECK(E.read_counters(&c));
    return c.h_capped > 0 ? SPHX_CAPPED : SPHX_OK;
}

extern "C" int sphx_div_curlVel(sphx_ctx *ctx, int mode, int64_t n, const double *pos,
                                const double *v, const double *rho, const double *m,
                                const double *h, int64_t row_beg, int64_t row_end, double *divV,
                                double *curlV, void *stream) {
    CTX_PROLOGUE(ctx);
    SPHX_CHECK(check_mode(mode));
    SPHX_CHECK(check_rows(n, row_beg, row_end));
    if (n == 0 || row_end == row_beg) return SPHX_OK;
    if (!pos || !v || !rho || !m || !h || !divV || !curlV) { sphx_set_error("null array"); return SPHX_ERR_ARG; }
    Engine &E = ctx->E;
    if (mode == SPHX_FP64) {
        SPHX_CHECK(ctx_load(ctx, n, pos, m, h, v, false, st));
        return plain_divcurl(E.s.x, E.s.y, E.s.z, E.s.vx, E.s.vy, E.s.vz, rho, E.s.m, E.s.h, n,
                             row_beg, row_end, divV, curlV, st);
    }
    SPHX_CHECK(ctx_load(ctx, n, pos, m, h, v, true, st));
    SPHX_CHECK(gather_in(E, rho, E.s.rho, st));
    SPHX_CHECK(E.build_boxes(st, true));
    SPHX_CHECK(E.divcurl(st));
    SPHX_CHECK(scatter_out(E, E.s.divv, row_beg, row_end, divV, st));
    SPHX_CHECK(scatter_out(E, E.s.curlv, row_beg, row_end, curlV, st));
    return SPHX_OK;
}

extern "C" int sphx_getAcc_sph(sphx_ctx *ctx, int mode, int64_t n, const double *pos,
                               const double *v, const double *rho, const double *P,
                               const double *c, const double *h, const double *m,
                               const double *divV, const double *curlV, const sphx_params *p,
                               int64_t row_beg, int64_t row_end, double *acc, double *dudt,
                               void *stream) {
    CTX_PROLOGUE(ctx);
    SPHX_CHECK(check_mode(mode));
    SPHX_CHECK(check_rows(n, row_beg, row_end));
    SPHX_CHECK(check_params(p));
    if (n == 0 || row_end == row_beg) return SPHX_OK;
    if (!pos || !v || !rho || !P || !c || !h || !m) { sphx_set_error("null array"); return SPHX_ERR_ARG; }
    if (p->balsara && (!divV || !curlV)) { sphx_set_error("Balsara switch needs divV and curlV"); return SPHX_ERR_ARG; }
    Engine &E = ctx->E;
    E.p = *p;
    bool sort = mode == SPHX_MIXED;
    SPHX_CHECK(ctx_load(ctx, n, pos, m, h, v, sort, st));
    SPHX_CHECK(gather_in(E, rho, E.s.rho, st));
    SPHX_CHECK(gather_in(E, P, E.s.P, st));
    SPHX_CHECK(gather_in(E, c, E.s.c, st));
    if (p->balsara) {
        SPHX_CHECK(gather_in(E, divV, E.s.divv, st));
        SPHX_CHECK(gather_in(E, curlV, E.s.curlv, st));
    }
    if (mode == SPHX_FP64) {
        SPHX_CHECK(plain_acc_sph(E.s, *p, n, 0, n, E.s.ax, E.s.ay, E.s.az, st));
        SPHX_CHECK(plain_du(E.s, *p, n, 0, n, E.s.du, st));
    } else {
        SPHX_CHECK(E.build_boxes(st, true));
        SPHX_CHECK(E.hydro(st));
    }
    SPHX_CHECK(scatter_out_aos(E, E.s.ax, E.s.ay, E.s.az, row_beg, row_end, acc, st));
    SPHX_CHECK(scatter_out(E, E.s.du, row_beg, row_end, dudt, st));
    return SPHX_OK;
}

extern "C" int sphx_getAcc_g_smth(sphx_ctx *ctx, int mode, int64_t n, const double *pos,
                                  const double *m, const double *eps, double G, int64_t row_beg,
                                  int64_t row_end, double *acc, double *pot, void *stream) {
    CTX_PROLOGUE(ctx);
    SPHX_CHECK(check_mode(mode));
    SPHX_CHECK(check_rows(n, row_beg, row_end));
    if (n == 0 || row_end == row_beg) return SPHX_OK;
    if (!pos || !m || !eps || !acc) { sphx_set_error("null array"); return SPHX_ERR_ARG; }
    Engine &E = ctx->E;
    E.p.G = G;
    E.p.eps_from_h = 0;
    if (mode == SPHX_FP64) {
        SPHX_CHECK(ctx_load(ctx, n, pos, m, nullptr, nullptr, false, st));
        SPHX_CHECK(plain_gravity(E.s.x, E.s.y, E.s.z, E.s.m, eps, G, n, row_beg, row_end, E.s.gx,
                                 E.s.gy, E.s.gz, pot, st));
        return sphx_soa_to_aos(E.s.gx, E.s.gy, E.s.gz, row_end - row_beg, acc, st);
    }
    SPHX_CHECK(ctx_load(ctx, n, pos, m, nullptr, nullptr, false, st));
    k_copy_d<<<gu(n), UT, 0, st>>>(eps, E.s.eps, n);
    k_copy_d<<<gu(n), UT, 0, st>>>(eps, E.s.h, n);
    SPHX_HIP(hipGetLastError());
    SPHX_CHECK(E.sort_and_permute(st, true));
    SPHX_CHECK(E.build_boxes(st, true));
    SPHX_CHECK(sphx_gravity_run(E, E.gcfg, pot != nullptr, st));
    SPHX_CHECK(scatter_out_aos(E, E.s.gx, E.s.gy, E.s.gz, row_beg, row_end, acc, st));
    if (pot) SPHX_CHECK(scatter_out(E, E.s.pot, row_beg, row_end, pot, st));
    return SPHX_OK;
}

extern "C" int sphx_pair_issue_rate(sphx_ctx *ctx, int32_t iters, double *pairs_per_s,
                                    void *stream) {
    CTX_PROLOGUE(ctx);
    if (!pairs_per_s || iters < 1 || iters > (1 << 16)) {
        sphx_set_error("pairs_per_s must be non-null and iters in 1..65536");
        return SPHX_ERR_ARG;
    }
    int cus = 0;
    SPHX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->E.device));
    // 32 workgroups of 256 lanes per CU: eight waves per SIMD, four rounds
    const int64_t blocks = (int64_t)cus * 32;
    float *out = nullptr;
    SPHX_HIP(hipMalloc(&out, sizeof(float) * blocks * 256));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    int rc = SPHX_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        sphx_set_error("hipEventCreate failed");
        rc = SPHX_ERR_HIP;
    }
    if (rc == SPHX_OK) rc = sphx_pair_probe_launch(blocks, iters, out, st);   // warm-up
    // the fastest of three timed launches
    for (int rep = 0; rep < 3 && rc == SPHX_OK; ++rep) {
        float t = 0.f;
        if (hipEventRecord(e0, st) != hipSuccess) {
            sphx_set_error("hipEventRecord failed");
            rc = SPHX_ERR_HIP;
        }
        if (rc == SPHX_OK) rc = sphx_pair_probe_launch(blocks, iters, out, st);
        if (rc == SPHX_OK && (hipEventRecord(e1, st) != hipSuccess ||
                              hipEventSynchronize(e1) != hipSuccess ||
                              hipEventElapsedTime(&t, e0, e1) != hipSuccess)) {
            sphx_set_error("timing the pair-rate reference failed");
            rc = SPHX_ERR_HIP;
        }
        if (rc == SPHX_OK && (rep == 0 || t < ms)) ms = t;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipStreamSynchronize(st);
    hipFree(out);
    if (rc != SPHX_OK) return rc;
    if (!(ms > 0.f)) {
        sphx_set_error("the pair-rate reference reported no elapsed time");
        return SPHX_ERR_HIP;
    }
    *pairs_per_s = (double)blocks * 256.0 * iters * 16.0 / (ms * 1e-3);
    return SPHX_OK;
}

extern "C" int sphx_neighbors(sphx_ctx *ctx, int64_t n, const double *pos, const double *h,
                              int64_t row_beg, int64_t row_end, int64_t *offsets, int32_t *idx,
                              int64_t capacity, int64_t *total, void *stream) {
    CTX_PROLOGUE(ctx);
    SPHX_CHECK(check_rows(n, row_beg, row_end));
    if (!offsets || !total) { sphx_set_error("null array"); return SPHX_ERR_ARG; }
    int64_t rows = row_end - row_beg;
    *total = 0;
    if (n == 0 || rows == 0) {
        int64_t z = 0;
        SPHX_HIP(hipMemcpy(offsets, &z, sizeof(z), hipMemcpyHostToDevice));
        return SPHX_OK;
    }
    Engine &E = ctx->E;
    SPHX_CHECK(ctx_load(ctx, n, pos, nullptr, h, nullptr, true, st));
    SPHX_CHECK(E.build_boxes(st, true));
    SPHX_CHECK(sphx_neighbor_pass(E, row_beg, row_end, E.tmpi, nullptr, nullptr, st));
    std::vector<int32_t> cnt(rows);
    SPHX_HIP(hipMemcpyAsync(cnt.data(), E.tmpi, sizeof(int32_t) * rows, hipMemcpyDeviceToHost, st));
    SPHX_HIP(hipStreamSynchronize(st));
    std::vector<int64_t> off(rows + 1, 0);
    for (int64_t r = 0; r < rows; ++r) off[r + 1] = off[r] + cnt[r];
    SPHX_HIP(hipMemcpy(offsets, off.data(), sizeof(int64_t) * (rows + 1), hipMemcpyHostToDevice));
    *total = off[rows];
    if (!idx || off[rows] > capacity) return SPHX_OK;
    SPHX_CHECK(sphx_neighbor_pass(E, row_beg, row_end, nullptr, offsets, idx, st));
    SPHX_HIP(hipStreamSynchronize(st));
    return SPHX_OK;
}

// Validate a cooling table (device arrays) on the host: sizes, units and
// strictly increasing positive nodes.
static int check_grid(const sphx_cool_grid *g, std::vector<double> *un, std::vector<double> *rn,
                      hipStream_t st) {
    if (!g || g->nu < 2 || g->nrho < 2 || !g->u_nodes || !g->rho_nodes || !g->du) {
        sphx_set_error("cooling table needs >= 2 nodes per axis and device arrays");
        return SPHX_ERR_ARG;
    }
    if (!(g->dt_ref_s > 0.0 && g->unit_u_cgs > 0.0 && g->unit_rho_cgs > 0.0 && g->unit_time_s > 0.0)) {
        sphx_set_error("cooling table needs positive dt_ref_s and unit factors");
        return SPHX_ERR_ARG;
    }
    un->resize(g->nu);
    rn->resize(g->nrho);
    SPHX_HIP(hipMemcpyAsync(un->data(), g->u_nodes, sizeof(double) * g->nu, hipMemcpyDeviceToHost, st));
    SPHX_HIP(hipMemcpyAsync(rn->data(), g->rho_nodes, sizeof(double) * g->nrho, hipMemcpyDeviceToHost, st));
    SPHX_HIP(hipStreamSynchronize(st));
    for (const std::vector<double> *v : {un, rn}) {
        if (!((*v)[0] > 0.0)) { sphx_set_error("cooling table nodes must be positive"); return SPHX_ERR_ARG; }
        for (size_t k = 1; k < v->size(); ++k)
            if (!((*v)[k] > (*v)[k - 1])) {
                sphx_set_error("cooling table nodes must increase strictly");
                return SPHX_ERR_ARG;
            }
    }
    return SPHX_OK;
}

static GridDev grid_view(const sphx_cool_grid *g) {
    GridDev d;
    d.nu = g->nu;
    d.nrho = g->nrho;
    d.u_nodes = g->u_nodes;
    d.rho_nodes = g->rho_nodes;
    d.du = g->du;
    d.dt_ref_s = g->dt_ref_s;
    d.unit_u_cgs = g->unit_u_cgs;
    d.unit_rho_cgs = g->unit_rho_cgs;
    d.unit_time_s = g->unit_time_s;
    return d;
}

// Copy a cooling table into engine-owned memory (kept until the engine dies).
int Engine::set_cool_grid(const sphx_cool_grid *g, hipStream_t st) {
    std::vector<double> un, rn;
    SPHX_CHECK(check_grid(g, &un, &rn, st));
    int64_t need = g->nu + g->nrho + g->nu * g->nrho;
    if (need > grid_cap) {
        if (grid_buf) hipFree(grid_buf);
        grid_buf = nullptr;
        grid_cap = 0;
        SPHX_HIP(hipMalloc(&grid_buf, sizeof(double) * need));
        grid_cap = need;
    }
    SPHX_HIP(hipMemcpyAsync(grid_buf, g->u_nodes, sizeof(double) * g->nu, hipMemcpyDeviceToDevice, st));
    SPHX_HIP(hipMemcpyAsync(grid_buf + g->nu, g->rho_nodes, sizeof(double) * g->nrho,
                            hipMemcpyDeviceToDevice, st));
    SPHX_HIP(hipMemcpyAsync(grid_buf + g->nu + g->nrho, g->du, sizeof(double) * g->nu * g->nrho,
                            hipMemcpyDeviceToDevice, st));
    SPHX_HIP(hipStreamSynchronize(st));
    grid = grid_view(g);
    grid.u_nodes = grid_buf;
    grid.rho_nodes = grid_buf + g->nu;
    grid.du = grid_buf + g->nu + g->nrho;
    drop_graph();
    return SPHX_OK;
}

// Validate a CHIMES table (device arrays) on the host: sizes, constants and
// strictly increasing axes (so that bisection finds the reference's node).
int sphx_check_chimes(const sphx_chimes_table *t, hipStream_t st) {
    if (!t || t->n_kpc < 1 || t->n_nh < 1 || t->n_z < 1 || t->n_t < 1 || t->n_m < 0 ||
        t->n_time < 1 || !t->kpc || !t->nh || !t->time_s || !t->u_evol || !t->mu ||
        (t->n_m > 0 && !t->metalz)) {
        sphx_set_error("CHIMES table needs positive sizes and device arrays");
        return SPHX_ERR_ARG;
    }
    if (t->z_index < 0 || t->z_index >= t->n_z) { sphx_set_error("CHIMES z_index outside the Z axis"); return SPHX_ERR_ARG; }
    if (!(t->xh > 0 && t->m_h > 0 && t->kpc_cm > 0 && t->unit_length_cm > 0 &&
          t->unit_rho_cgs > 0 && t->unit_u_cgs > 0 && t->unit_time_s > 0) ||
        std::isnan(t->u_max_cgs)) {
        sphx_set_error("CHIMES table needs positive constants and unit factors");
        return SPHX_ERR_ARG;
    }
    const float *axes[3] = {t->kpc, t->nh, t->time_s};
    int32_t lens[3] = {t->n_kpc, t->n_nh, t->n_time};
    for (int a = 0; a < 3; ++a) {
        std::vector<float> v(lens[a]);
        SPHX_HIP(hipMemcpyAsync(v.data(), axes[a], sizeof(float) * lens[a], hipMemcpyDeviceToHost, st));
        SPHX_HIP(hipStreamSynchronize(st));
        for (int k = 0; k < lens[a]; ++k)
            if (std::isnan(v[k]) || (k > 0 && !(v[k] > v[k - 1]))) {
                sphx_set_error("CHIMES table axes (kpc, nH, time) must increase strictly");
                return SPHX_ERR_ARG;
            }
    }
    return SPHX_OK;
}

ChimesDev sphx_chimes_view(const sphx_chimes_table *t) {
    ChimesDev c;
    c.n_kpc = t->n_kpc;
    c.n_nh = t->n_nh;
    c.n_z = t->n_z;
    c.n_t = t->n_t;
    c.n_m = t->n_m;
    c.n_time = t->n_time;
    c.z_index = t->z_index;
    c.kpc = t->kpc;
    c.nh = t->nh;
    c.time_s = t->time_s;
    c.u_evol = t->u_evol;
    c.mu = t->mu;
    c.metalz = t->metalz;
    c.xh = t->xh;
    c.m_h = t->m_h;
    c.u_max_cgs = t->u_max_cgs;
    c.kpc_cm = t->kpc_cm;
    c.unit_length_cm = t->unit_length_cm;
    c.unit_rho_cgs = t->unit_rho_cgs;
    c.unit_u_cgs = t->unit_u_cgs;
    c.unit_time_s = t->unit_time_s;
    return c;
}

// Copy a CHIMES table into engine-owned memory; keep_ions also sizes the
// per-particle species fractions (cap x n_m, zeroed).
int Engine::set_chimes(const sphx_chimes_table *t, bool keep_ions, hipStream_t st) {
    SPHX_CHECK(sphx_check_chimes(t, st));
    int64_t cells = (int64_t)t->n_kpc * t->n_t * t->n_nh * t->n_z * t->n_time;
    int64_t need = t->n_kpc + t->n_nh + t->n_time + 2 * cells + cells * t->n_m;
    if (need > chimes_cap) {
        if (chimes_buf) hipFree(chimes_buf);
        chimes_buf = nullptr;
        chimes_cap = 0;
        SPHX_HIP(hipMalloc(&chimes_buf, sizeof(float) * need));
        chimes_cap = need;
    }
    float *q = chimes_buf;
    ChimesDev c = sphx_chimes_view(t);
    const float *src[6] = {t->kpc, t->nh, t->time_s, t->u_evol, t->mu, t->metalz};
    int64_t len[6] = {t->n_kpc, t->n_nh, t->n_time, cells, cells, cells * t->n_m};
    const float **dst[6] = {&c.kpc, &c.nh, &c.time_s, &c.u_evol, &c.mu, &c.metalz};
    for (int a = 0; a < 6; ++a) {
        if (len[a] > 0)
            SPHX_HIP(hipMemcpyAsync(q, src[a], sizeof(float) * len[a], hipMemcpyDeviceToDevice, st));
        *dst[a] = q;
        q += len[a];
    }
    c.ion = nullptr;
    if (keep_ions && t->n_m > 0) {
        int64_t ion_need = cap * t->n_m;
        if (ion_need > chimes_ion_cap) {
            if (chimes_ion) hipFree(chimes_ion);
            chimes_ion = nullptr;
            chimes_ion_cap = 0;
            SPHX_HIP(hipMalloc(&chimes_ion, sizeof(float) * ion_need));
            chimes_ion_cap = ion_need;
        }
        SPHX_HIP(hipMemsetAsync(chimes_ion, 0, sizeof(float) * ion_need, st));
        c.ion = chimes_ion;
    }
    SPHX_HIP(hipStreamSynchronize(st));
    chimes = c;
    drop_graph();
    return SPHX_OK;
}

int sphx_copy_cool(Engine &E, const sphx_cooling *cool) {
    if (!cool) {
        E.has_cool = false;
        return SPHX_OK;
    }
    if (cool->ntab < 2 || !cool->T || !cool->L || !(cool->unit_u_cgs > 0) ||
        !(cool->unit_rho_cgs > 0) || !(cool->unit_time_s > 0) || !(cool->tol > 0) ||
        cool->max_iter <= 0) {
        sphx_set_error("invalid sphx_cooling");
        return SPHX_ERR_ARG;
    }
    for (int k = 1; k < cool->ntab; ++k)
        if (!(cool->T[k] > cool->T[k - 1])) {
            sphx_set_error("cooling table T must increase");
            return SPHX_ERR_ARG;
        }
    if (E.cool_T) hipFree(E.cool_T);
    if (E.cool_L) hipFree(E.cool_L);
    E.cool_T = E.cool_L = nullptr;
    SPHX_HIP(hipMalloc(&E.cool_T, sizeof(double) * cool->ntab));
    SPHX_HIP(hipMalloc(&E.cool_L, sizeof(double) * cool->ntab));
    SPHX_HIP(hipMemcpy(E.cool_T, cool->T, sizeof(double) * cool->ntab, hipMemcpyHostToDevice));
    SPHX_HIP(hipMemcpy(E.cool_L, cool->L, sizeof(double) * cool->ntab, hipMemcpyHostToDevice));
    E.cool = *cool;
    E.cool.T = nullptr;
    E.cool.L = nullptr;
    E.has_cool = true;
    return SPHX_OK;
}

extern "C" int sphx_cool(sphx_ctx *ctx, int64_t n, const double *rho, const double *u, double dt,
                         const sphx_cooling *cool, double *u_out, int32_t *iters, void *stream) {
    CTX_PROLOGUE(ctx);
    if (n < 0 || (n > 0 && (!rho || !u || !u_out))) { sphx_set_error("bad arguments"); return SPHX_ERR_ARG; }
    Engine &E = ctx->E;
    SPHX_CHECK(E.reserve(64));
    SPHX_CHECK(sphx_copy_cool(E, cool));
    if (!E.has_cool) { sphx_set_error("cooling table required"); return SPHX_ERR_ARG; }
    if (n == 0) return SPHX_OK;
    SPHX_CHECK(sphx_zero_counters(E.counters, st));
    SPHX_CHECK(sphx_cool_launch(sphx_engine_cool(E), rho, u, dt, nullptr, nullptr, nullptr, n,
                                u_out, iters, E.counters, st));
    SPHX_HIP(hipStreamSynchronize(st));
    sphx_counters c{};
    SPHX_CHECK(E.read_counters(&c));
    return c.cool_capped > 0 ? SPHX_CAPPED : SPHX_OK;
}

extern "C" int sphx_cool_network(sphx_ctx *ctx, int64_t n, const double *rho, const double *u,
                                 double dt, const sphx_network *net, double *u_out,
                                 int32_t *iters, void *stream) {
    CTX_PROLOGUE(ctx);
    if (n < 0 || (n > 0 && (!rho || !u || !u_out))) { sphx_set_error("bad arguments"); return SPHX_ERR_ARG; }
    if (!(dt >= 0.0)) { sphx_set_error("dt must be >= 0"); return SPHX_ERR_ARG; }
    NetDev nd;
    SPHX_CHECK(sphx_network_dev(net, &nd));
    Engine &E = ctx->E;
    SPHX_CHECK(E.reserve(64));
    if (n == 0) return SPHX_OK;
    SPHX_CHECK(sphx_zero_counters(E.counters, st));
    SPHX_CHECK(sphx_net_launch(nd, rho, u, dt, n, u_out, iters, E.counters, st));
    SPHX_HIP(hipStreamSynchronize(st));
    sphx_counters c{};
    SPHX_CHECK(E.read_counters(&c));
    return c.cool_capped > 0 ? SPHX_CAPPED : SPHX_OK;
}

extern "C" int sphx_cool_grid_apply(sphx_ctx *ctx, int64_t n, const double *rho,
                                    const double *u_in, double dt, const sphx_cool_grid *grid,
                                    double *u_out, void *stream) {
    CTX_PROLOGUE(ctx);
    if (n < 0 || (n > 0 && (!rho || !u_in || !u_out))) { sphx_set_error("bad arguments"); return SPHX_ERR_ARG; }
    std::vector<double> un, rn;
    SPHX_CHECK(check_grid(grid, &un, &rn, st));
    return sphx_grid_apply_launch(grid_view(grid), rho, u_in, dt, n, u_out, st);
}

extern "C" int sphx_chimes_apply(sphx_ctx *ctx, int64_t n, const double *pos, const double *rho,
                                 const double *u_in, double dt, const sphx_chimes_table *table,
                                 double *u_out, float *ion_out, double *mu_out, void *stream) {
    CTX_PROLOGUE(ctx);
    if (n < 0 || (n > 0 && (!pos || !rho || !u_in || !u_out))) { sphx_set_error("bad arguments"); return SPHX_ERR_ARG; }
    if (std::isnan(dt)) { sphx_set_error("dt is NaN"); return SPHX_ERR_ARG; }
    SPHX_CHECK(sphx_check_chimes(table, st));
    if (ion_out && table->n_m == 0) ion_out = nullptr;
    SPHX_CHECK(sphx_chimes_apply_launch(sphx_chimes_view(table), pos, rho, u_in, dt, n, u_out,
                                        ion_out, mu_out, st));
    SPHX_HIP(hipStreamSynchronize(st));
    return SPHX_OK;
}

extern "C" int sphx_cool_grid_build(sphx_ctx *ctx, const sphx_cooling *cool,
                                    const sphx_cool_grid *grid, void *stream) {
    CTX_PROLOGUE(ctx);
    Engine &E = ctx->E;
    std::vector<double> un, rn;
    SPHX_CHECK(check_grid(grid, &un, &rn, st));
    SPHX_CHECK(E.reserve(64));
    SPHX_CHECK(sphx_copy_cool(E, cool));
    if (!E.has_cool) { sphx_set_error("cooling table required"); return SPHX_ERR_ARG; }
    SPHX_CHECK(sphx_grid_build_launch(grid_view(grid), sphx_engine_cool(E), grid->du, st));
    SPHX_HIP(hipStreamSynchronize(st));
    return SPHX_OK;
}

extern "C" int sphx_diagnostics(sphx_ctx *ctx, int64_t n, const double *pos, const double *v,
                                const double *m, const double *u, const double *eps, double G,
                                sphx_diag *out, void *stream) {
    CTX_PROLOGUE(ctx);
    if (n < 0 || !out || (n > 0 && (!v || !m))) { sphx_set_error("bad arguments"); return SPHX_ERR_ARG; }
    if (n == 0) {  // no particles: every sum is zero
        *out = sphx_diag{};
        return SPHX_OK;
    }
    Engine &E = ctx->E;
    SPHX_CHECK(ctx_load(ctx, n, pos ? pos : v, m, nullptr, v, false, st));
    const double *pot = nullptr;
    if (pos && eps && n > 0) {
        SPHX_CHECK(plain_gravity(E.s.x, E.s.y, E.s.z, E.s.m, eps, G, n, 0, n, E.s.gx, E.s.gy,
                                 E.s.gz, E.s.pot, st));
        pot = E.s.pot;
    }
    return sphx_diag_reduce(pos ? E.s.x : nullptr, E.s.y, E.s.z, E.s.vx, E.s.vy, E.s.vz, E.s.m, u,
                            pot, n, E.tmpd, out, st);
}

extern "C" int sphx_morton_sort(sphx_ctx *ctx, int64_t n, const double *pos, uint64_t *keys_out,
                                int32_t *perm_out, void *stream) {
    CTX_PROLOGUE(ctx);
    if (n < 0 || (n > 0 && (!pos || !keys_out || !perm_out))) { sphx_set_error("bad arguments"); return SPHX_ERR_ARG; }
    if (n == 0) return SPHX_OK;
    Engine &E = ctx->E;
    SPHX_CHECK(ctx_load(ctx, n, pos, nullptr, nullptr, nullptr, false, st));
    SPHX_CHECK(E.sort_and_permute(st, false));
    SPHX_HIP(hipMemcpyAsync(keys_out, E.keys, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, st));
    SPHX_HIP(hipMemcpyAsync(perm_out, E.perm, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
    SPHX_HIP(hipStreamSynchronize(st));
    return SPHX_OK;
}

// ---------------------------------------------------------------------------
// Resident plan
// ---------------------------------------------------------------------------
struct sphx_plan {
    Engine E;
    bool loaded = false;
    bool inited = false;
    // adaptive time step: configuration, and whether the next sphx_plan_step's
    // dt argument seeds the first adaptive step
    int32_t adaptive = 0;
    double c_cfl = 0.25, dt_min = 1e-5, dt_max = 1e-3;
    bool seed_pending = false;
    // caller-owned state (sphx_plan_bind): read at the start of every init /
    // step call, written back at its end, in the caller's particle order
    bool bound = false;
    double *bpos = nullptr, *bvel = nullptr, *bu = nullptr, *bh = nullptr;
};

// Bound buffers -> the plan's working arrays (engine order).
static int plan_pull_bound(sphx_plan *plan, hipStream_t st) {
    Engine &E = plan->E;
    if (!plan->bound || E.n == 0) return SPHX_OK;
    k_gather_aos_by_id<<<gu(E.n), UT, 0, st>>>(plan->bpos, E.s.id, E.s.x, E.s.y, E.s.z, E.n);
    k_gather_aos_by_id<<<gu(E.n), UT, 0, st>>>(plan->bvel, E.s.id, E.s.vx, E.s.vy, E.s.vz, E.n);
    k_gather_by_id<<<gu(E.n), UT, 0, st>>>(plan->bu, E.s.id, E.s.u, E.n);
    k_gather_by_id<<<gu(E.n), UT, 0, st>>>(plan->bh, E.s.id, E.s.h, E.n);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

// The plan's working arrays -> bound buffers (caller order).
static int plan_push_bound(sphx_plan *plan, hipStream_t st) {
    Engine &E = plan->E;
    if (!plan->bound || E.n == 0) return SPHX_OK;
    SPHX_CHECK(scatter_out_aos(E, E.s.x, E.s.y, E.s.z, 0, E.n, plan->bpos, st));
    SPHX_CHECK(scatter_out_aos(E, E.s.vx, E.s.vy, E.s.vz, 0, E.n, plan->bvel, st));
    SPHX_CHECK(scatter_out(E, E.s.u, 0, E.n, plan->bu, st));
    SPHX_CHECK(scatter_out(E, E.s.h, 0, E.n, plan->bh, st));
    return SPHX_OK;
}

#define PLAN_PROLOGUE(plan)                        \
    if (!plan) {                                   \
        sphx_set_error("null plan");               \
        return SPHX_ERR_ARG;                       \
    }                                              \
    DeviceGuard guard_(plan->E.device);            \
    hipStream_t st = (hipStream_t)stream;

extern "C" int sphx_plan_create(int device, int mode, const sphx_params *p,
                                const sphx_cooling *cool, int64_t capacity, sphx_plan **out) {
    if (!out) { sphx_set_error("null output pointer"); return SPHX_ERR_ARG; }
    SPHX_CHECK(check_mode(mode));
    SPHX_CHECK(check_params(p));
    if (mode == SPHX_FP64 && p->h_mode == 1) {
        sphx_set_error("kNN smoothing lengths (h_mode 1) are a mixed-precision option");
        return SPHX_ERR_ARG;
    }
    if (capacity <= 0 || capacity > INT32_MAX) { sphx_set_error("bad capacity"); return SPHX_ERR_ARG; }
    int ndev = 0;
    SPHX_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) { sphx_set_error("no such HIP device"); return SPHX_ERR_ARG; }
    DeviceGuard guard_(device);
    sphx_plan *pl = new sphx_plan();
    pl->E.device = device;
    pl->E.mode = mode;
    pl->E.p = *p;
    pl->E.want_lists = true;
    int rc = pl->E.reserve(std::max<int64_t>(capacity, 64));
    if (rc >= 0) rc = sphx_copy_cool(pl->E, cool);
    if (rc < 0) {
        delete pl;
        return rc;
    }
    *out = pl;
    return SPHX_OK;
}

extern "C" int sphx_plan_load(sphx_plan *plan, int64_t n, const double *pos, const double *vel,
                              const double *u, const double *m, const double *h, const double *eps,
                              void *stream) {
    PLAN_PROLOGUE(plan);
    Engine &E = plan->E;
    if (n <= 0 || n > E.cap) { sphx_set_error("particle count must be 1 .. the plan capacity"); return SPHX_ERR_ARG; }
    if (!pos || !vel || !u || !m || !h) { sphx_set_error("null array"); return SPHX_ERR_ARG; }
    if (!E.p.eps_from_h && !eps) { sphx_set_error("eps required when eps_from_h == 0"); return SPHX_ERR_ARG; }
    SPHX_CHECK(E.set_runs({n}, st));
    SPHX_CHECK(sphx_aos_to_soa(pos, n, E.s.x, E.s.y, E.s.z, st));
    SPHX_CHECK(sphx_aos_to_soa(vel, n, E.s.vx, E.s.vy, E.s.vz, st));
    k_copy_d<<<gu(n), UT, 0, st>>>(u, E.s.u, n);
    k_copy_d<<<gu(n), UT, 0, st>>>(m, E.s.m, n);
    k_copy_d<<<gu(n), UT, 0, st>>>(h, E.s.h, n);
    k_copy_d<<<gu(n), UT, 0, st>>>(eps ? eps : h, E.s.eps, n);
    k_fill_d<<<gu(n), UT, 0, st>>>(E.s.ut_prev, 0.0, n);
    k_fill_d<<<gu(n), UT, 0, st>>>(E.s.divv, 0.0, n);
    k_fill_d<<<gu(n), UT, 0, st>>>(E.s.ax, 0.0, n);
    k_fill_d<<<gu(n), UT, 0, st>>>(E.s.ay, 0.0, n);
    k_fill_d<<<gu(n), UT, 0, st>>>(E.s.az, 0.0, n);
    k_fill_i<<<gu(n), UT, 0, st>>>(E.s.k, E.p.nngb, n);
    k_iota<<<gu(n), UT, 0, st>>>(E.s.id, n);
    k_fill_i<<<gu(n), UT, 0, st>>>(E.s.run, 0, n);
    SPHX_HIP(hipGetLastError());
    E.steps = 0;
    plan->loaded = true;
    plan->inited = false;
    plan->bound = false;
    return SPHX_OK;
}

extern "C" int sphx_plan_bind(sphx_plan *plan, int64_t n, double *pos, double *vel, double *u,
                              const double *m, double *h, const double *eps, void *stream) {
    SPHX_CHECK(sphx_plan_load(plan, n, pos, vel, u, m, h, eps, stream));
    plan->bound = true;
    plan->bpos = pos;
    plan->bvel = vel;
    plan->bu = u;
    plan->bh = h;
    return SPHX_OK;
}

extern "C" int sphx_plan_init(sphx_plan *plan, void *stream) {
    PLAN_PROLOGUE(plan);
    if (!plan->loaded) { sphx_set_error("plan has no particles loaded"); return SPHX_ERR_STATE; }
    Engine &E = plan->E;
    RunParams rp0{0.0, 0, E.p.eps_from_h};
    rp0.adaptive = plan->adaptive;
    rp0.c_cfl = plan->c_cfl;
    rp0.dt_min = plan->dt_min;
    rp0.dt_max = plan->dt_max;
    plan->seed_pending = plan->adaptive != 0;
    E.run_params.assign(1, rp0);
    SPHX_CHECK(E.upload_run_params(st));
    SPHX_CHECK(sphx_zero_counters(E.counters, st));
    SPHX_CHECK(plan_pull_bound(plan, st));
    if (E.mode == SPHX_MIXED) {
        SPHX_CHECK(E.sort_and_permute(st, true));
        SPHX_CHECK(E.forces(st, false, false));
    } else {
        SPHX_CHECK(E.plain_forces(st, false, false));
    }
    SPHX_CHECK(E.bootstrap_finish(st));
    SPHX_CHECK(plan_push_bound(plan, st));
    E.blk_started = false;
    E.blk_tick = 0;
    plan->inited = true;
    return SPHX_OK;
}

extern "C" int sphx_plan_step(sphx_plan *plan, int32_t n_steps, double dt, int32_t use_cooling,
                              void *stream) {
    PLAN_PROLOGUE(plan);
    if (!plan->inited) { sphx_set_error("call sphx_plan_init first"); return SPHX_ERR_STATE; }
    if (n_steps < 0 || !(dt > 0.0)) { sphx_set_error("bad step count or dt"); return SPHX_ERR_ARG; }
    Engine &E = plan->E;
    if (use_cooling < 0 || use_cooling > 4) { sphx_set_error("use_cooling is 0, 1 (solve), 2 (2-D table), 3 (CHIMES table) or 4 (H/He network)"); return SPHX_ERR_ARG; }
    if (use_cooling == 4 && !E.net.on) { sphx_set_error("plan has no cooling network"); return SPHX_ERR_ARG; }
    if (use_cooling == 1 && !E.has_cool) { sphx_set_error("plan has no cooling table"); return SPHX_ERR_ARG; }
    if (use_cooling == 2 && !E.grid.du) { sphx_set_error("plan has no 2-D cooling table"); return SPHX_ERR_ARG; }
    if (use_cooling == 3 && !E.chimes.u_evol) { sphx_set_error("plan has no CHIMES table"); return SPHX_ERR_ARG; }
    // The host copy of the run parameters is current except for t and, with
    // the adaptive step, dt: pull the device copy before changing anything.
    RunParams &hp = E.run_params[0];
    int32_t uc = use_cooling;
    // block time steps: dt is the longest step, the run advances by substeps
    double run_dt = dt;
    if (E.blk_B > 0) {
        if (E.blk_started && dt != E.blk_dtmax) {
            sphx_set_error("block time steps keep the longest step fixed; call sphx_plan_init to change it");
            return SPHX_ERR_ARG;
        }
        E.blk_dtmax = dt;
        run_dt = std::ldexp(dt, -E.blk_B);
    }
    bool set_dt = plan->adaptive ? plan->seed_pending : hp.dt != run_dt;
    if (set_dt || hp.use_cool != uc) {
        SPHX_CHECK(E.pull_run_params(st));
        if (set_dt) E.run_params[0].dt = run_dt;
        E.run_params[0].use_cool = uc;
        plan->seed_pending = false;
        SPHX_CHECK(E.upload_run_params(st));
    }
    // cooling is gated by the run's use_cool flag (device memory), so a
    // captured step graph stays valid when the flag changes
    // the Barnes-Hut build sizes its levels on the host, so those steps run eagerly
    SPHX_CHECK(plan_pull_bound(plan, st));
    SPHX_CHECK((E.use_graph && !E.timing && E.bh_theta == 0.0 && E.blk_B == 0)
                   ? E.step_graphs(st, n_steps)
                   : E.step(st, n_steps));
    return plan_push_bound(plan, st);
}

extern "C" int sphx_plan_set_adaptive_dt(sphx_plan *plan, int32_t on, double c_cfl,
                                         double dt_min, double dt_max) {
    void *stream = nullptr;
    PLAN_PROLOGUE(plan);
    if (on && !(c_cfl > 0.0 && dt_min > 0.0 && dt_max >= dt_min)) {
        sphx_set_error("adaptive dt needs c_cfl > 0 and 0 < dt_min <= dt_max");
        return SPHX_ERR_ARG;
    }
    if (on && plan->E.blk_B > 0) {
        sphx_set_error("the adaptive global dt and block time steps exclude each other");
        return SPHX_ERR_ARG;
    }
    plan->adaptive = on ? 1 : 0;
    if (on) {
        plan->c_cfl = c_cfl;
        plan->dt_min = dt_min;
        plan->dt_max = dt_max;
    }
    plan->seed_pending = plan->adaptive != 0;
    if (plan->inited) {
        Engine &E = plan->E;
        SPHX_CHECK(E.pull_run_params(st));
        RunParams &q = E.run_params[0];
        q.adaptive = plan->adaptive;
        q.c_cfl = plan->c_cfl;
        q.dt_min = plan->dt_min;
        q.dt_max = plan->dt_max;
        SPHX_CHECK(E.upload_run_params(st));
        SPHX_HIP(hipStreamSynchronize(st));
    }
    return SPHX_OK;
}

extern "C" int sphx_plan_set_block_steps(sphx_plan *plan, int32_t levels, double c_cfl) {
    if (!plan) { sphx_set_error("null plan"); return SPHX_ERR_ARG; }
    Engine &E = plan->E;
    if (levels < 0 || levels > 20) { sphx_set_error("block levels must be 0..20"); return SPHX_ERR_ARG; }
    if (levels > 0 && E.mode != SPHX_MIXED) { sphx_set_error("block time steps are a mixed-precision option"); return SPHX_ERR_ARG; }
    if (levels > 0 && plan->adaptive) { sphx_set_error("block time steps replace the adaptive global dt; switch that off first"); return SPHX_ERR_ARG; }
    if (levels > 0 && !(c_cfl > 0.0)) { sphx_set_error("c_cfl must be positive"); return SPHX_ERR_ARG; }
    if (E.blk_started) { sphx_set_error("the run is inside block steps; call sphx_plan_init first"); return SPHX_ERR_STATE; }
    E.blk_B = levels;
    E.blk_c_cfl = c_cfl;
    E.blk_tick = 0;
    E.drop_graph();
    return SPHX_OK;
}

extern "C" int sphx_plan_block_bins(sphx_plan *plan, int32_t *out, void *stream) {
    PLAN_PROLOGUE(plan);
    Engine &E = plan->E;
    if (!plan->inited || E.blk_B == 0 || !E.blk_started) {
        sphx_set_error("no block steps taken since sphx_plan_init");
        return SPHX_ERR_STATE;
    }
    if (!out) { sphx_set_error("null output"); return SPHX_ERR_ARG; }
    SPHX_CHECK(scatter_out_i(E, E.s.bin, 0, E.n, out, st));
    SPHX_HIP(hipStreamSynchronize(st));
    return SPHX_OK;
}

extern "C" int sphx_plan_set_gravity_bh(sphx_plan *plan, double theta) {
    if (!plan) { sphx_set_error("null plan"); return SPHX_ERR_ARG; }
    if (theta < 0.0 || theta > 2.0) { sphx_set_error("theta must be 0 (direct sum) or in (0, 2]"); return SPHX_ERR_ARG; }
    if (theta > 0.0 && plan->E.mode != SPHX_MIXED) {
        sphx_set_error("Barnes-Hut gravity is a mixed-precision option");
        return SPHX_ERR_ARG;
    }
    plan->E.bh_theta = theta;
    plan->E.drop_graph();
    return SPHX_OK;
}

static void ckpt_header(CkptHeader &h, const Engine &E) {
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, SPHX_CKPT_MAGIC, 8);
    h.version = SPHX_CKPT_VERSION;
    h.mode = E.mode;
    h.p = E.p;
    h.g = E.gcfg;
    h.bh_theta = E.bh_theta;
    h.morton_bits = E.morton_bits();
}

extern "C" int sphx_plan_save(sphx_plan *plan, const char *path, void *stream) {
    PLAN_PROLOGUE(plan);
    if (!plan->inited) { sphx_set_error("nothing to save: call sphx_plan_init first"); return SPHX_ERR_STATE; }
    if (plan->E.blk_B > 0) { sphx_set_error("checkpoints of block-step runs are not supported"); return SPHX_ERR_STATE; }
    Engine &E = plan->E;
    SPHX_CHECK(E.pull_run_params(st));
    CkptHeader h;
    ckpt_header(h, E);
    h.n = E.n;
    h.steps = E.steps;
    h.rp = E.run_params[0];
    h.seed_pending = plan->seed_pending ? 1 : 0;
    return sphx_ckpt_write(path, h, E, 0, st);
}

extern "C" int sphx_plan_restore(sphx_plan *plan, const char *path, void *stream) {
    PLAN_PROLOGUE(plan);
    Engine &E = plan->E;
    CkptHeader h;
    SPHX_CHECK(sphx_ckpt_read_header(path, &h));
    if (h.mode != E.mode) { sphx_set_error("checkpoint precision mode differs from the plan's"); return SPHX_ERR_ARG; }
    if (h.n <= 0 || h.n > E.cap) { sphx_set_error("checkpoint particle count exceeds the plan capacity"); return SPHX_ERR_ARG; }
    if ((h.rp.use_cool == 1 && !E.has_cool) || (h.rp.use_cool == 2 && !E.grid.du) ||
        (h.rp.use_cool == 3 && !E.chimes.u_evol) || (h.rp.use_cool == 4 && !E.net.on)) {
        sphx_set_error("checkpoint run cools with a table the plan does not have");
        return SPHX_ERR_ARG;
    }
    if (h.morton_bits != E.morton_bits()) { sphx_set_error("checkpoint from an engine with another key layout"); return SPHX_ERR_ARG; }
    E.p = h.p;
    SPHX_CHECK(E.set_runs({h.n}, st));
    SPHX_CHECK(sphx_ckpt_upload(path, E, 0, h.n, st));
    k_fill_i<<<gu(h.n), UT, 0, st>>>(E.s.run, 0, h.n);
    SPHX_HIP(hipGetLastError());
    E.run_params.assign(1, h.rp);
    SPHX_CHECK(E.upload_run_params(st));
    E.steps = (int)h.steps;
    E.gcfg = h.g;
    E.bh_theta = h.bh_theta;
    plan->adaptive = h.rp.adaptive;
    if (h.rp.adaptive) {
        plan->c_cfl = h.rp.c_cfl;
        plan->dt_min = h.rp.dt_min;
        plan->dt_max = h.rp.dt_max;
    }
    plan->seed_pending = h.seed_pending != 0;
    plan->loaded = plan->inited = true;
    plan->bound = false;  // the restored run replaces a bound one
    E.drop_graph();
    SPHX_HIP(hipStreamSynchronize(st));
    return SPHX_OK;
}

extern "C" int sphx_plan_set_cooling_grid(sphx_plan *plan, const sphx_cool_grid *grid) {
    void *stream = nullptr;
    PLAN_PROLOGUE(plan);
    return plan->E.set_cool_grid(grid, st);
}

extern "C" int sphx_plan_set_chimes_table(sphx_plan *plan, const sphx_chimes_table *table) {
    void *stream = nullptr;
    PLAN_PROLOGUE(plan);
    return plan->E.set_chimes(table, true, st);
}

extern "C" int sphx_plan_set_cooling_network(sphx_plan *plan, const sphx_network *net) {
    if (!plan) { sphx_set_error("null plan"); return SPHX_ERR_ARG; }
    NetDev nd;
    SPHX_CHECK(sphx_network_dev(net, &nd));
    plan->E.net = nd;
    plan->E.drop_graph();  // the network's parameters are captured by value
    return SPHX_OK;
}

extern "C" int sphx_plan_chimes_ions(sphx_plan *plan, float *out, void *stream) {
    PLAN_PROLOGUE(plan);
    Engine &E = plan->E;
    if (!E.chimes.ion) { sphx_set_error("plan has no CHIMES table with species"); return SPHX_ERR_STATE; }
    if (!out && E.n > 0) { sphx_set_error("null output"); return SPHX_ERR_ARG; }
    if (E.n > 0)
        SPHX_HIP(hipMemcpyAsync(out, E.chimes.ion, sizeof(float) * E.n * E.chimes.n_m,
                                hipMemcpyDeviceToDevice, st));
    SPHX_HIP(hipStreamSynchronize(st));
    return SPHX_OK;
}

extern "C" int sphx_plan_time(sphx_plan *plan, double *t, double *dt, double *dt_parts,
                              void *stream) {
    PLAN_PROLOGUE(plan);
    if (!plan->inited) { sphx_set_error("call sphx_plan_init first"); return SPHX_ERR_STATE; }
    Engine &E = plan->E;
    SPHX_CHECK(E.pull_run_params(st));
    const RunParams &q = E.run_params[0];
    if (t) *t = q.t;
    if (dt) *dt = q.dt;
    if (dt_parts)
        for (int k = 0; k < 4; ++k) dt_parts[k] = q.dt_parts[k];
    return SPHX_OK;
}

extern "C" int sphx_global_dt(sphx_ctx *ctx, int64_t n, const double *h, const double *c,
                              const double *acc, const double *acc_g, const double *divV,
                              double c_cfl, double dt_min, double dt_max, double *out,
                              void *stream) {
    CTX_PROLOGUE(ctx);
    if (n <= 0 || !h || !c || !acc || !acc_g || !divV || !out) {
        sphx_set_error("bad arguments");
        return SPHX_ERR_ARG;
    }
    if (!(c_cfl > 0.0 && dt_min > 0.0 && dt_max >= dt_min)) {
        sphx_set_error("global dt needs c_cfl > 0 and 0 < dt_min <= dt_max");
        return SPHX_ERR_ARG;
    }
    Engine &E = ctx->E;
    SPHX_CHECK(E.reserve(std::max<int64_t>(n, 4096)));
    return sphx_global_dt_launch(h, c, acc, 3, acc_g, 3, divV, n, c_cfl, dt_min, dt_max, E.tmpd,
                                 E.cap, out, st);
}

extern "C" int sphx_plan_store(sphx_plan *plan, double *pos, double *vel, double *u, double *h,
                               double *rho, double *acc, double *acc_g, double *dudt, int32_t *k,
                               void *stream) {
    PLAN_PROLOGUE(plan);
    if (!plan->loaded) { sphx_set_error("plan has no particles loaded"); return SPHX_ERR_STATE; }
    Engine &E = plan->E;
    int64_t n = E.n;
    SPHX_CHECK(scatter_out_aos(E, E.s.x, E.s.y, E.s.z, 0, n, pos, st));
    SPHX_CHECK(scatter_out_aos(E, E.s.vx, E.s.vy, E.s.vz, 0, n, vel, st));
    SPHX_CHECK(scatter_out(E, E.s.u, 0, n, u, st));
    SPHX_CHECK(scatter_out(E, E.s.h, 0, n, h, st));
    SPHX_CHECK(scatter_out(E, E.s.rho, 0, n, rho, st));
    SPHX_CHECK(scatter_out_aos(E, E.s.ax, E.s.ay, E.s.az, 0, n, acc, st));
    SPHX_CHECK(scatter_out_aos(E, E.s.gx, E.s.gy, E.s.gz, 0, n, acc_g, st));
    SPHX_CHECK(scatter_out(E, E.s.du, 0, n, dudt, st));
    SPHX_CHECK(scatter_out_i(E, E.s.k, 0, n, k, st));
    return SPHX_OK;
}

extern "C" int sphx_plan_query(sphx_plan *plan, sphx_counters *out) {
    if (!plan || !out) { sphx_set_error("null argument"); return SPHX_ERR_ARG; }
    DeviceGuard guard_(plan->E.device);
    SPHX_HIP(hipDeviceSynchronize());
    return plan->E.read_counters(out);
}

extern "C" int sphx_plan_diagnostics(sphx_plan *plan, int32_t with_pe, sphx_diag *out,
                                     void *stream) {
    PLAN_PROLOGUE(plan);
    if (!out) { sphx_set_error("null argument"); return SPHX_ERR_ARG; }
    if (!plan->loaded) { sphx_set_error("plan has no particles loaded"); return SPHX_ERR_STATE; }
    return plan->E.diagnostics(st, with_pe != 0, out);
}

extern "C" int sphx_plan_destroy(sphx_plan *plan) {
    if (!plan) return SPHX_OK;
    delete plan;
    return SPHX_OK;
}

// Gravity launch configuration of a plan (used by the autotuner).
// Tiled sum: ipl 1/2/4 particles per lane, j tiles of 256/512/1024, js j
// splits (1..64).  Symmetric sum: ipl 8, tj 512, js = block pairs per wave
// (1..4096).
static bool gravity_config_ok(int32_t ipl, int32_t tj, int32_t js) {
    // symmetric sum: js block pairs per wave, or (js = 1) a block pair split
    // into 512 / tj parts of tj j particles
    if (ipl == 8) return (tj == 512 && js >= 1 && js <= 4096) || ((tj == 256 || tj == 128) && js == 1);
    return (ipl == 1 || ipl == 2 || ipl == 4) && (tj == 256 || tj == 512 || tj == 1024) &&
           js >= 1 && js <= 64 && !(ipl == 1 && tj == 1024) && !(ipl == 4 && tj == 1024);
}

extern "C" int sphx_plan_set_gravity_config(sphx_plan *plan, int32_t ipl, int32_t tj, int32_t js) {
    if (!plan) { sphx_set_error("null plan"); return SPHX_ERR_ARG; }
    if (!gravity_config_ok(ipl, tj, js)) {
        sphx_set_error("unsupported gravity configuration");
        return SPHX_ERR_ARG;
    }
    plan->E.gcfg = GravConfig{ipl, tj, js};
    plan->E.drop_graph();  // a captured step holds the previous launch shape
    return SPHX_OK;
}

extern "C" int sphx_ctx_set_gravity_config(sphx_ctx *ctx, int32_t ipl, int32_t tj, int32_t js) {
    if (!ctx) { sphx_set_error("null context"); return SPHX_ERR_ARG; }
    if (!gravity_config_ok(ipl, tj, js)) {
        sphx_set_error("unsupported gravity configuration");
        return SPHX_ERR_ARG;
    }
    ctx->E.gcfg = GravConfig{ipl, tj, js};
    return SPHX_OK;
}

int sphx_set_cell_level(Engine &E, int32_t bias) {
    if (bias < -3 || bias > 3) { sphx_set_error("cell level bias must be in -3..3"); return SPHX_ERR_ARG; }
    E.lcut_bias = bias;
    if (E.R > 0) SPHX_CHECK(E.set_runs(E.run_n));  // rebuild the run tables with the new cut levels
    return SPHX_OK;
}

extern "C" int sphx_plan_set_cell_level(sphx_plan *plan, int32_t bias) {
    void *stream = nullptr;
    PLAN_PROLOGUE(plan);
    (void)st;
    return sphx_set_cell_level(plan->E, bias);
}

extern "C" int sphx_plan_set_neighbor_lists(sphx_plan *plan, int32_t on) {
    if (!plan) { sphx_set_error("null plan"); return SPHX_ERR_ARG; }
    if ((on != 0) != plan->E.use_lists) {
        plan->E.use_lists = on != 0;
        plan->E.drop_graph();  // the captured steps launch the other kernels
    }
    return SPHX_OK;
}

extern "C" int sphx_plan_set_timing(sphx_plan *plan, int32_t on) {
    if (!plan) { sphx_set_error("null plan"); return SPHX_ERR_ARG; }
    plan->E.timing = on != 0;
    for (int k = 0; k < PH_COUNT; ++k) {
        plan->E.phase_ms[k] = 0.0;
        plan->E.phase_n[k] = 0;
    }
    return SPHX_OK;
}

extern "C" int sphx_plan_timing(sphx_plan *plan, double *ms, int64_t *count, int32_t nphase) {
    if (!plan || !ms || !count || nphase < PH_COUNT) { sphx_set_error("bad arguments"); return SPHX_ERR_ARG; }
    for (int k = 0; k < PH_COUNT; ++k) {
        ms[k] = plan->E.phase_ms[k];
        count[k] = plan->E.phase_n[k];
    }
    return SPHX_OK;
}

extern "C" int sphx_plan_set_graph(sphx_plan *plan, int32_t on) {
    if (!plan) { sphx_set_error("null plan"); return SPHX_ERR_ARG; }
    plan->E.use_graph = on != 0;
    plan->E.drop_graph();
    return SPHX_OK;
}

extern "C" int sphx_bh_gravity(sphx_ctx *ctx, int64_t n, const double *pos, const double *m,
                               const double *eps, double G, double theta, int64_t row_beg,
                               int64_t row_end, double *acc, double *pot, int64_t *stats,
                               void *stream) {
    CTX_PROLOGUE(ctx);
    SPHX_CHECK(check_rows(n, row_beg, row_end));
    if (n == 0 || row_end == row_beg) return SPHX_OK;
    if (!pos || !m || !eps || !acc) { sphx_set_error("null array"); return SPHX_ERR_ARG; }
    if (!(theta > 0.0) || theta > 2.0) { sphx_set_error("theta must be in (0, 2]"); return SPHX_ERR_ARG; }
    Engine &E = ctx->E;
    E.p.G = G;
    E.p.eps_from_h = 0;
    E.max_runs = 1;
    SPHX_CHECK(ctx_load(ctx, n, pos, m, nullptr, nullptr, false, st));
    k_copy_d<<<gu(n), UT, 0, st>>>(eps, E.s.eps, n);
    k_copy_d<<<gu(n), UT, 0, st>>>(eps, E.s.h, n);
    SPHX_HIP(hipGetLastError());
    SPHX_CHECK(E.sort_and_permute(st, true));
    if (!ctx->bh) ctx->bh = sphx_bh_host_new();
    BhStats bs{};
    SPHX_CHECK(sphx_bh_run(E, *ctx->bh, theta, pot != nullptr, st, &bs));
    SPHX_CHECK(scatter_out_aos(E, E.s.gx, E.s.gy, E.s.gz, row_beg, row_end, acc, st));
    if (pot) SPHX_CHECK(scatter_out(E, E.s.pot, row_beg, row_end, pot, st));
    if (stats) {
        stats[0] = bs.nodes;
        stats[1] = bs.depth;
        stats[2] = bs.interactions;
        stats[3] = bs.overflow;
        stats[4] = bs.max_fill;
        stats[5] = (int64_t)(bs.build_ms * 1000.0);
        stats[6] = (int64_t)(bs.walk_ms * 1000.0);
        stats[7] = bs.groups;
    }
    SPHX_HIP(hipStreamSynchronize(st));
    if (bs.overflow > 0) {
        sphx_set_error("Barnes-Hut frontier overflow: increase theta");
        return SPHX_ERR_STATE;
    }
    return SPHX_OK;
}
