// Served runs: many independent simulations packed into one engine.
//
// The reference runs one simulation per process (its drivers own all
// allocations, hfv_sph_gpu_v2_evrard.cu:150-189).  Here independent runs of
// uneven N share one packed particle array: each run is a contiguous,
// Morton-sorted segment with its own buckets, supers, gravity i-blocks and
// j-tiles, so one launch per pass advances every live run and no pass lets
// one run see another.  Admission appends and bootstraps only the new run;
// retirement copies the run out in its caller order and compacts the arrays.
#include <algorithm>

#include <cstring>

#include "sphx_internal.h"

int sphx_zero_counters(int64_t *c, hipStream_t st);

namespace {

constexpr int BT = 256;
inline unsigned gb(int64_t n) { return (unsigned)((n + BT - 1) / BT); }

__global__ void kb_iota(int32_t *a, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i < n) a[i] = (int32_t)i;
}
__global__ void kb_fill_i(int32_t *a, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void kb_fill_d(double *a, double v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void kb_copy_d(const double *__restrict__ a, double *__restrict__ b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i < n) b[i] = a[i];
}
__global__ void kb_copy_i(const int32_t *__restrict__ a, int32_t *__restrict__ b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i < n) b[i] = a[i];
}
__global__ void kb_add_i(int32_t *a, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i < n) a[i] += v;
}
// dst[i] = src[perm[i] - base] over one segment (perm holds absolute indices)
__global__ void kb_gather_d(const double *__restrict__ src, const int32_t *__restrict__ perm,
                            int32_t base, double *__restrict__ dst, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i < n) dst[i] = src[perm[i] - base];
}
__global__ void kb_gather_i(const int32_t *__restrict__ src, const int32_t *__restrict__ perm,
                            int32_t base, int32_t *__restrict__ dst, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i < n) dst[i] = src[perm[i] - base];
}
__global__ void kb_aos_to_soa(const double *__restrict__ a, int64_t n, double *x, double *y,
                              double *z) {
    int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i >= n) return;
    x[i] = a[3 * i]; y[i] = a[3 * i + 1]; z[i] = a[3 * i + 2];
}
// out[id[i]] = v[i] for a run segment (id is run-local)
__global__ void kb_scatter_d(const double *__restrict__ v, const int32_t *__restrict__ id,
                             double *__restrict__ out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i < n) out[id[i]] = v[i];
}
__global__ void kb_scatter_i(const int32_t *__restrict__ v, const int32_t *__restrict__ id,
                             int32_t *__restrict__ out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i < n) out[id[i]] = v[i];
}
__global__ void kb_scatter_aos(const double *__restrict__ a, const double *__restrict__ b,
                               const double *__restrict__ c, const int32_t *__restrict__ id,
                               double *__restrict__ out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i >= n) return;
    int64_t o = 3 * (int64_t)id[i];
    out[o] = a[i]; out[o + 1] = b[i]; out[o + 2] = c[i];
}

}  // namespace

int sphx_launch_morton(Engine &E, hipStream_t st);
int sphx_copy_cool(Engine &E, const sphx_cooling *cool);

struct BatchRun {
    int64_t id;
    int64_t n;
    int64_t steps;
};

struct sphx_batch {
    Engine E;
    std::vector<BatchRun> runs;  // packed order
    int64_t next_id = 1;
};

// Arrays that travel with a particle between steps: the persistent state, the
// total acceleration (needed by the next half kick) and the density (returned
// when the run retires).
static std::vector<double *> moving_d(Engine &E) {
    Soa &s = E.s;
    return {s.x, s.y, s.z, s.vx, s.vy, s.vz, s.u, s.m, s.h, s.eps, s.ut_prev, s.divv,
            s.ax, s.ay, s.az, s.rho};
}
static std::vector<int32_t *> moving_i(Engine &E) { return {E.s.k, E.s.id, E.s.run}; }

// Morton-sort one run's segment [p0, p1) in place (other runs untouched).
static int sort_segment(Engine &E, int r, hipStream_t st) {
    int64_t p0 = E.hr_pbeg[r], p1 = E.hr_pend[r], n = p1 - p0;
    if (n <= 1) return SPHX_OK;
    // keys for the whole packed set are cheap; only the segment is sorted
    SPHX_CHECK(sphx_launch_morton(E, st));
    int rb = 0;
    while ((1 << rb) < E.max_runs) ++rb;
    int bits = 3 * sphx_morton_bits(E.max_runs) + rb;
    SPHX_CHECK(sphx_radix_sort(E.keys + p0, E.perm + p0, E.keys2 + p0, E.perm2 + p0, E.hist, n,
                               bits, st));
    for (double *a : moving_d(E)) {
        kb_gather_d<<<gb(n), BT, 0, st>>>(a + p0, E.perm + p0, (int32_t)p0, E.tmpd, n);
        kb_copy_d<<<gb(n), BT, 0, st>>>(E.tmpd, a + p0, n);
    }
    for (int32_t *a : moving_i(E)) {
        kb_gather_i<<<gb(n), BT, 0, st>>>(a + p0, E.perm + p0, (int32_t)p0, E.tmpi, n);
        kb_copy_i<<<gb(n), BT, 0, st>>>(E.tmpi, a + p0, n);
    }
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

#define BATCH_PROLOGUE(b)                                \
    if (!b) {                                            \
        sphx_set_error("null batch");                    \
        return SPHX_ERR_ARG;                             \
    }                                                    \
    int prev_dev_ = 0;                                   \
    hipGetDevice(&prev_dev_);                            \
    hipSetDevice(b->E.device);                           \
    struct Restore_ {                                    \
        int d;                                           \
        ~Restore_() { hipSetDevice(d); }                 \
    } restore_{prev_dev_};                               \
    hipStream_t st = (hipStream_t)stream;

extern "C" int sphx_batch_create(int device, int mode, const sphx_params *p,
                                 const sphx_cooling *cool, int64_t capacity, int32_t max_runs,
                                 sphx_batch **out) {
    if (!out || !p) { sphx_set_error("null argument"); return SPHX_ERR_ARG; }
    if (mode != SPHX_MIXED) { sphx_set_error("served runs use SPHX_MIXED"); return SPHX_ERR_ARG; }
    if (capacity <= 0 || capacity > INT32_MAX || max_runs < 1 || max_runs > 4096) {
        sphx_set_error("bad capacity or max_runs");
        return SPHX_ERR_ARG;
    }
    int ndev = 0;
    SPHX_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) { sphx_set_error("no such HIP device"); return SPHX_ERR_ARG; }
    int prev = 0;
    hipGetDevice(&prev);
    hipSetDevice(device);
    sphx_batch *b = new sphx_batch();
    b->E.device = device;
    b->E.mode = SPHX_MIXED;
    b->E.p = *p;
    b->E.max_runs = max_runs;
    b->E.gcfg = GravConfig{2, 512, 1};  // fixed shape: a run's sums do not depend on the mix
    b->E.want_lists = true;
    int rc = b->E.reserve(std::max<int64_t>(capacity, 64));
    if (rc >= 0 && cool) rc = sphx_copy_cool(b->E, cool);
    if (rc >= 0) rc = b->E.set_runs({});
    hipSetDevice(prev);
    if (rc < 0) {
        delete b;
        return rc;
    }
    *out = b;
    return SPHX_OK;
}

extern "C" int sphx_batch_destroy(sphx_batch *b) {
    if (!b) return SPHX_OK;
    int prev = 0;
    hipGetDevice(&prev);
    hipSetDevice(b->E.device);
    delete b;
    hipSetDevice(prev);
    return SPHX_OK;
}

extern "C" int sphx_batch_admit(sphx_batch *b, int64_t n, const double *pos, const double *vel,
                                const double *u, const double *m, const double *h,
                                const double *eps, double dt, int32_t use_cooling,
                                int64_t *run_id, void *stream) {
    BATCH_PROLOGUE(b);
    Engine &E = b->E;
    if (n <= 0 || !pos || !vel || !u || !m || !h || !run_id || !(dt > 0.0)) {
        sphx_set_error("bad admission arguments");
        return SPHX_ERR_ARG;
    }
    if (!E.p.eps_from_h && !eps) { sphx_set_error("eps required when eps_from_h == 0"); return SPHX_ERR_ARG; }
    if (use_cooling < 0 || use_cooling > 4) { sphx_set_error("use_cooling is 0, 1 (solve), 2 (2-D table), 3 (CHIMES table) or 4 (H/He network)"); return SPHX_ERR_ARG; }
    if (use_cooling == 4 && !E.net.on) { sphx_set_error("batch has no cooling network"); return SPHX_ERR_ARG; }
    if (use_cooling == 1 && !E.has_cool) { sphx_set_error("batch has no cooling table"); return SPHX_ERR_ARG; }
    if (use_cooling == 2 && !E.grid.du) { sphx_set_error("batch has no 2-D cooling table"); return SPHX_ERR_ARG; }
    if (use_cooling == 3 && !E.chimes.u_evol) { sphx_set_error("batch has no CHIMES table"); return SPHX_ERR_ARG; }
    if (E.n + n > E.cap || (int)b->runs.size() >= E.max_runs) {
        sphx_set_error("admission refused: particle capacity or run slots exhausted");
        return SPHX_REFUSED;
    }
    if (E.gcfg.ipl == 8) {
        std::vector<int64_t> with_new = E.run_n;
        with_new.push_back(n);
        if (!sphx_sym_fits(E, with_new, E.gcfg.js)) {
            sphx_set_error("admission refused: a row of the run's gravity partial sums exceeds "
                           "the reserved memory");
            return SPHX_REFUSED;
        }
    }
    int64_t p0 = E.n;
    int r = (int)b->runs.size();
    kb_aos_to_soa<<<gb(n), BT, 0, st>>>(pos, n, E.s.x + p0, E.s.y + p0, E.s.z + p0);
    kb_aos_to_soa<<<gb(n), BT, 0, st>>>(vel, n, E.s.vx + p0, E.s.vy + p0, E.s.vz + p0);
    kb_copy_d<<<gb(n), BT, 0, st>>>(u, E.s.u + p0, n);
    kb_copy_d<<<gb(n), BT, 0, st>>>(m, E.s.m + p0, n);
    kb_copy_d<<<gb(n), BT, 0, st>>>(h, E.s.h + p0, n);
    kb_copy_d<<<gb(n), BT, 0, st>>>(eps ? eps : h, E.s.eps + p0, n);
    for (double *a : {E.s.ut_prev, E.s.divv, E.s.ax, E.s.ay, E.s.az})
        kb_fill_d<<<gb(n), BT, 0, st>>>(a + p0, 0.0, n);
    kb_fill_i<<<gb(n), BT, 0, st>>>(E.s.k + p0, E.p.nngb, n);
    kb_iota<<<gb(n), BT, 0, st>>>(E.s.id + p0, n);
    kb_fill_i<<<gb(n), BT, 0, st>>>(E.s.run + p0, r, n);
    SPHX_HIP(hipGetLastError());
    SPHX_CHECK(E.pull_run_params(st));   // device dt (adaptive runs) and t
    std::vector<int64_t> sizes = E.run_n;
    sizes.push_back(n);
    SPHX_CHECK(E.set_runs(sizes, st));
    E.run_params.resize(E.R);
    E.run_params[r] = RunParams{dt, use_cooling, E.p.eps_from_h};
    SPHX_CHECK(E.upload_run_params(st));
    E.drop_graph();
    // bootstrap only the new run: sort its segment, forces at t = 0, du_prev
    SPHX_CHECK(sort_segment(E, r, st));
    SPHX_CHECK(sphx_zero_counters(E.counters, st));
    SPHX_CHECK(E.restrict_to(r));
    int rc = E.forces(st, false, false);
    if (rc >= 0) rc = E.bootstrap_finish(st);
    E.restrict_to(-1);
    SPHX_CHECK(rc);
    BatchRun br{b->next_id++, n, 0};
    b->runs.push_back(br);
    *run_id = br.id;
    return SPHX_OK;
}

static int find_run(sphx_batch *b, int64_t id) {
    for (size_t r = 0; r < b->runs.size(); ++r)
        if (b->runs[r].id == id) return (int)r;
    return -1;
}

static int store_run(sphx_batch *b, int r, double *pos, double *vel, double *u, double *h,
                     double *rho, double *acc, int32_t *k, hipStream_t st) {
    Engine &E = b->E;
    int64_t p0 = E.hr_pbeg[r], p1 = E.hr_pend[r], n = p1 - p0;
    const int32_t *id = E.s.id + p0;
    if (pos) kb_scatter_aos<<<gb(n), BT, 0, st>>>(E.s.x + p0, E.s.y + p0, E.s.z + p0, id, pos, n);
    if (vel) kb_scatter_aos<<<gb(n), BT, 0, st>>>(E.s.vx + p0, E.s.vy + p0, E.s.vz + p0, id, vel, n);
    if (acc) kb_scatter_aos<<<gb(n), BT, 0, st>>>(E.s.ax + p0, E.s.ay + p0, E.s.az + p0, id, acc, n);
    if (u) kb_scatter_d<<<gb(n), BT, 0, st>>>(E.s.u + p0, id, u, n);
    if (h) kb_scatter_d<<<gb(n), BT, 0, st>>>(E.s.h + p0, id, h, n);
    if (rho) kb_scatter_d<<<gb(n), BT, 0, st>>>(E.s.rho + p0, id, rho, n);
    if (k) kb_scatter_i<<<gb(n), BT, 0, st>>>(E.s.k + p0, id, k, n);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

extern "C" int sphx_batch_store(sphx_batch *b, int64_t run_id, double *pos, double *vel,
                                double *u, double *h, double *rho, double *acc, int32_t *k,
                                void *stream) {
    BATCH_PROLOGUE(b);
    int r = find_run(b, run_id);
    if (r < 0) { sphx_set_error("no such run"); return SPHX_ERR_ARG; }
    return store_run(b, r, pos, vel, u, h, rho, acc, k, st);
}

extern "C" int sphx_batch_retire(sphx_batch *b, int64_t run_id, double *pos, double *vel,
                                 double *u, double *h, double *rho, double *acc, int32_t *k,
                                 void *stream) {
    BATCH_PROLOGUE(b);
    Engine &E = b->E;
    int r = find_run(b, run_id);
    if (r < 0) { sphx_set_error("no such run"); return SPHX_ERR_ARG; }
    SPHX_CHECK(store_run(b, r, pos, vel, u, h, rho, acc, k, st));
    int64_t p0 = E.hr_pbeg[r], p1 = E.hr_pend[r], n = p1 - p0;
    (void)n;
    // compact: move the later runs down by n (through scratch, the ranges overlap)
    int64_t tail = E.n - p1;
    if (tail > 0) {
        for (double *a : moving_d(E)) {
            kb_copy_d<<<gb(tail), BT, 0, st>>>(a + p1, E.tmpd, tail);
            kb_copy_d<<<gb(tail), BT, 0, st>>>(E.tmpd, a + p0, tail);
        }
        for (int32_t *a : moving_i(E)) {
            kb_copy_i<<<gb(tail), BT, 0, st>>>(a + p1, E.tmpi, tail);
            kb_copy_i<<<gb(tail), BT, 0, st>>>(E.tmpi, a + p0, tail);
        }
        kb_add_i<<<gb(tail), BT, 0, st>>>(E.s.run + p0, -1, tail);
        SPHX_HIP(hipGetLastError());
    }
    SPHX_CHECK(E.pull_run_params(st));
    std::vector<int64_t> sizes = E.run_n;
    sizes.erase(sizes.begin() + r);
    E.run_params.erase(E.run_params.begin() + r);
    b->runs.erase(b->runs.begin() + r);
    SPHX_CHECK(E.set_runs(sizes, st));
    SPHX_CHECK(E.upload_run_params(st));
    E.drop_graph();
    return SPHX_OK;
}

extern "C" int sphx_batch_step(sphx_batch *b, int32_t n_steps, void *stream) {
    BATCH_PROLOGUE(b);
    Engine &E = b->E;
    if (n_steps < 0) { sphx_set_error("bad step count"); return SPHX_ERR_ARG; }
    if (b->runs.empty() || n_steps == 0) return SPHX_OK;
    // per-run use_cool flags (uploaded run params) gate the cooling solve
    if (E.use_graph) SPHX_CHECK(E.step_graphs(st, n_steps));
    else SPHX_CHECK(E.step(st, n_steps));
    for (auto &rr : b->runs) rr.steps += n_steps;
    return SPHX_OK;
}

extern "C" int sphx_batch_set_gravity_symmetric(sphx_batch *b, int32_t block_pairs,
                                                int64_t budget_bytes) {
    void *stream = nullptr;
    BATCH_PROLOGUE(b);
    (void)st;
    if (!b->runs.empty()) { sphx_set_error("set the gravity sum before admitting runs"); return SPHX_ERR_STATE; }
    if (block_pairs < 1 || block_pairs > 4096 || budget_bytes < 0) {
        sphx_set_error("block_pairs must be 1..4096 and the budget non-negative");
        return SPHX_ERR_ARG;
    }
    SPHX_CHECK(sphx_sym_reserve(b->E, block_pairs, budget_bytes));
    b->E.gcfg = GravConfig{8, 512, block_pairs};
    b->E.drop_graph();
    return SPHX_OK;
}

extern "C" int sphx_batch_set_cell_level(sphx_batch *b, int32_t bias) {
    void *stream = nullptr;
    BATCH_PROLOGUE(b);
    (void)st;
    return sphx_set_cell_level(b->E, bias);
}

extern "C" int64_t sphx_gravity_symmetric_bytes(int64_t n, int32_t block_pairs) {
    if (n <= 0 || block_pairs < 1) return 0;
    return sphx_sym_run_bytes(n, block_pairs);
}

extern "C" int64_t sphx_gravity_symmetric_min_bytes(int64_t n, int32_t block_pairs) {
    if (n <= 0 || block_pairs < 1) return 0;
    return sphx_sym_min_bytes(n, block_pairs);
}

extern "C" int sphx_batch_set_chimes_table(sphx_batch *b, const sphx_chimes_table *table) {
    void *stream = nullptr;
    BATCH_PROLOGUE(b);
    return b->E.set_chimes(table, false, st);
}

extern "C" int sphx_batch_set_cooling_network(sphx_batch *b, const sphx_network *net) {
    if (!b) { sphx_set_error("null batch"); return SPHX_ERR_ARG; }
    NetDev nd;
    SPHX_CHECK(sphx_network_dev(net, &nd));
    b->E.net = nd;
    b->E.drop_graph();
    return SPHX_OK;
}

extern "C" int sphx_batch_set_cooling_grid(sphx_batch *b, const sphx_cool_grid *grid) {
    void *stream = nullptr;
    BATCH_PROLOGUE(b);
    return b->E.set_cool_grid(grid, st);
}

extern "C" int sphx_batch_set_adaptive_dt(sphx_batch *b, int64_t run_id, int32_t on,
                                          double c_cfl, double dt_min, double dt_max,
                                          void *stream) {
    BATCH_PROLOGUE(b);
    Engine &E = b->E;
    int r = find_run(b, run_id);
    if (r < 0) { sphx_set_error("no such run"); return SPHX_ERR_ARG; }
    if (on && !(c_cfl > 0.0 && dt_min > 0.0 && dt_max >= dt_min)) {
        sphx_set_error("adaptive dt needs c_cfl > 0 and 0 < dt_min <= dt_max");
        return SPHX_ERR_ARG;
    }
    SPHX_CHECK(E.pull_run_params(st));
    RunParams &q = E.run_params[r];
    q.adaptive = on ? 1 : 0;
    if (on) {
        q.c_cfl = c_cfl;
        q.dt_min = dt_min;
        q.dt_max = dt_max;
    }
    SPHX_CHECK(E.upload_run_params(st));
    SPHX_HIP(hipStreamSynchronize(st));
    return SPHX_OK;
}

extern "C" int sphx_batch_time(sphx_batch *b, int64_t run_id, double *t, double *dt,
                               double *dt_parts, void *stream) {
    BATCH_PROLOGUE(b);
    Engine &E = b->E;
    int r = find_run(b, run_id);
    if (r < 0) { sphx_set_error("no such run"); return SPHX_ERR_ARG; }
    SPHX_CHECK(E.pull_run_params(st));
    const RunParams &q = E.run_params[r];
    if (t) *t = q.t;
    if (dt) *dt = q.dt;
    if (dt_parts)
        for (int k = 0; k < 4; ++k) dt_parts[k] = q.dt_parts[k];
    return SPHX_OK;
}

extern "C" int sphx_batch_save_run(sphx_batch *b, int64_t run_id, const char *path,
                                   void *stream) {
    BATCH_PROLOGUE(b);
    Engine &E = b->E;
    int r = find_run(b, run_id);
    if (r < 0) { sphx_set_error("no such run"); return SPHX_ERR_ARG; }
    SPHX_CHECK(E.pull_run_params(st));
    CkptHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, SPHX_CKPT_MAGIC, 8);
    h.version = SPHX_CKPT_VERSION;
    h.mode = E.mode;
    h.p = E.p;
    h.g = E.gcfg;
    h.morton_bits = E.morton_bits();
    h.n = E.hr_pend[r] - E.hr_pbeg[r];
    h.steps = b->runs[r].steps;
    h.rp = E.run_params[r];
    return sphx_ckpt_write(path, h, E, E.hr_pbeg[r], st);
}

// Re-admit a saved run: its particles are appended in their saved internal
// order with their saved state, so no sort and no bootstrap are needed.
extern "C" int sphx_batch_restore_run(sphx_batch *b, const char *path, int64_t *run_id,
                                      void *stream) {
    BATCH_PROLOGUE(b);
    Engine &E = b->E;
    if (!run_id) { sphx_set_error("null argument"); return SPHX_ERR_ARG; }
    CkptHeader h;
    SPHX_CHECK(sphx_ckpt_read_header(path, &h));
    if (h.mode != E.mode || std::memcmp(&h.p, &E.p, sizeof(sphx_params)) != 0) {
        sphx_set_error("checkpoint physics parameters differ from the batch's");
        return SPHX_ERR_ARG;
    }
    if (h.morton_bits != E.morton_bits()) {
        sphx_set_error("checkpoint from an engine with another key layout");
        return SPHX_ERR_ARG;
    }
    if ((h.rp.use_cool == 1 && !E.has_cool) || (h.rp.use_cool == 2 && !E.grid.du) ||
        (h.rp.use_cool == 3 && !E.chimes.u_evol) || (h.rp.use_cool == 4 && !E.net.on)) {
        sphx_set_error("checkpoint run cools with a table the batch does not have");
        return SPHX_ERR_ARG;
    }
    if (h.n <= 0) { sphx_set_error("empty checkpoint"); return SPHX_ERR_ARG; }
    if (E.n + h.n > E.cap || (int)b->runs.size() >= E.max_runs) {
        sphx_set_error("admission refused: particle capacity or run slots exhausted");
        return SPHX_REFUSED;
    }
    if (E.gcfg.ipl == 8) {
        std::vector<int64_t> with_new = E.run_n;
        with_new.push_back(h.n);
        if (!sphx_sym_fits(E, with_new, E.gcfg.js)) {
            sphx_set_error("admission refused: a row of the run's gravity partial sums exceeds "
                           "the reserved memory");
            return SPHX_REFUSED;
        }
    }
    int64_t p0 = E.n, n = h.n;
    int r = (int)b->runs.size();
    SPHX_CHECK(sphx_ckpt_upload(path, E, p0, n, st));
    kb_fill_i<<<gb(n), BT, 0, st>>>(E.s.run + p0, r, n);
    SPHX_HIP(hipGetLastError());
    SPHX_CHECK(E.pull_run_params(st));
    std::vector<int64_t> sizes = E.run_n;
    sizes.push_back(n);
    SPHX_CHECK(E.set_runs(sizes, st));
    E.run_params.resize(E.R);
    E.run_params[r] = h.rp;
    SPHX_CHECK(E.upload_run_params(st));
    E.drop_graph();
    BatchRun br{b->next_id++, n, h.steps};
    b->runs.push_back(br);
    *run_id = br.id;
    SPHX_HIP(hipStreamSynchronize(st));
    return SPHX_OK;
}

extern "C" int sphx_batch_runs(sphx_batch *b, int64_t *ids, int64_t *sizes, int64_t *steps,
                               int32_t capacity, int32_t *count) {
    if (!b || !count) { sphx_set_error("null argument"); return SPHX_ERR_ARG; }
    *count = (int32_t)b->runs.size();
    for (int32_t r = 0; r < *count && r < capacity; ++r) {
        if (ids) ids[r] = b->runs[r].id;
        if (sizes) sizes[r] = b->runs[r].n;
        if (steps) steps[r] = b->runs[r].steps;
    }
    return SPHX_OK;
}

extern "C" int sphx_batch_set_graph(sphx_batch *b, int32_t on) {
    if (!b) { sphx_set_error("null batch"); return SPHX_ERR_ARG; }
    b->E.use_graph = on != 0;
    b->E.drop_graph();
    return SPHX_OK;
}

extern "C" int sphx_batch_query(sphx_batch *b, sphx_counters *out) {
    if (!b || !out) { sphx_set_error("null argument"); return SPHX_ERR_ARG; }
    int prev = 0;
    hipGetDevice(&prev);
    hipSetDevice(b->E.device);
    hipDeviceSynchronize();
    int rc = b->E.read_counters(out);
    hipSetDevice(prev);
    return rc;
}

extern "C" int64_t sphx_batch_bytes_per_particle(void) {
    // 25 FP64 arrays + 7 int32 arrays + 2 uint64 keys + tables (< 1 B amortised)
    // + the neighbour lists of the pair passes (SPHX_NBR_MAX int32 entries, an
    // 80-byte record and a Balsara factor per particle)
    return 25 * 8 + 7 * 4 + 2 * 8 + 1 + SPHX_NBR_MAX * 4 + (int64_t)sizeof(PairRec) + 4;
}
