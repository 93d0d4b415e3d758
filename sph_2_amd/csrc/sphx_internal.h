// Internal declarations of libsphx: device helpers, the Engine that owns the
// packed particle state and its scratch, and the kernel launchers.
//
// Layout (HBM, all SoA, Morton order inside each run):
//   FP64 state  x y z vx vy vz u m h eps ut_prev ax ay az (total) gx gy gz
//               (gravity) rho P c divv curlv du pot
//   int32       k (neighbour count), id (caller's original index), run
//   buckets     up to 64 consecutive particles of one run (one wavefront);
//               FP64 AABB + max h + max eps per bucket
//   supers      up to 64 consecutive buckets of one run, FP64 AABB + max h
//   candidates  per bucket, the buckets of its run within reach, CSR int32
// Runs are contiguous segments of the packed arrays; a single simulation is
// one run.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/sphx.h"

#define SPHX_BUCKET 64
// keys per workgroup tile of the radix sort (its histograms: 256 per tile)
#define SPHX_RS_TILE 1024
#define SPHX_SUPER 64
#define SPHX_NBR_MAX 128  // entries of a particle's neighbour list (more: its bucket walks)

// What the pair passes stage for a neighbour j, written by the density pass
// for the neighbour-list passes: FP64 position, h and velocity, and the FP32
// values the walks derive from the FP64 state (m, rho, P / rho^2, c), plus the
// length of j's own list.  80 bytes, read as five 16-byte loads.
struct alignas(16) PairRec {
    double x, y, z, h;
    double vx, vy, vz;
    float m, rho;
    float pr, c;
    int32_t len, pad;
};
static_assert(sizeof(PairRec) == 80, "PairRec layout");

// ---------------------------------------------------------------------------
// Error plumbing
// ---------------------------------------------------------------------------
void sphx_set_error(const std::string &msg);

#define SPHX_HIP(call)                                                                 \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            sphx_set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
            return SPHX_ERR_HIP;                                                       \
        }                                                                              \
    } while (0)

#define SPHX_CHECK(rc)                 \
    do {                               \
        int rc_ = (rc);                \
        if (rc_ < 0) return rc_;       \
    } while (0)

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
#define SPHX_SIGMA_D (0.31830988618379067154)  // 1/pi
#define SPHX_SIGMA_F (0.31830988618379067154f)

// Pair distance^2 exactly as the oracle computes it: separately rounded
// products, summed left to right.  Neighbour decisions depend on it.
__device__ __forceinline__ double sphx_r2(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    double a = dx * dx;
    double b = dy * dy;
    double c = dz * dz;
    return (a + b) + c;
}

__device__ __forceinline__ double sphx_sq(double t) {
#pragma clang fp contract(off)
    return t * t;
}

// Step of the reference's smoothing-length rules as a fraction of h_prev:
// (coeff * 2) with coeff = 0.001 (h_mode 2: LIBSX3:952, DRV2:286), 0.01
// (h_mode 3: do_smoothingX4) or, at the start, 0.005 (h_mode 4: the LIBV7
// driver, hfvGpuSphTyp_v7.cu:28); 2 x fl(coeff) is fl(2 coeff) exactly.
inline double sphx_h_step(int h_mode) {
    return h_mode == 3 ? 0.02 : (h_mode == 4 ? 0.01 : 0.002);
}
// Pass limit of the h_mode 4 rule (N_iter > 500, LIBV7:321).
constexpr int SPHX_H4_PASSES = 500;
// Pass limit of the h_mode 5 rule (h_smooth_fast, LIBSX:1133-1184), which has
// none in the reference: a row whose count jumps across the band at a 0.3 %
// step would loop there for ever.  A row that converges moves h by at most
// a factor 1.003^10000 (about e^30) first.
constexpr int SPHX_H5_PASSES = 10000;

// cbrt(nngb/k) clamped to [1/2, 2] by a fixed-count Newton iteration with
// separately rounded operations; identical on any IEEE machine.
__device__ __forceinline__ double sphx_h_factor(int nngb, int k) {
#pragma clang fp contract(off)
    if (8 * (long long)k <= (long long)nngb) return 2.0;
    if ((long long)k >= 8 * (long long)nngb) return 0.5;
    double x = (double)nngb / (double)k;
    double y = 1.0;
    for (int it = 0; it < 12; ++it) {
        double y2 = y * y;
        double y3 = y2 * y;
        double num = y3 - x;
        double den = 3.0 * y2;
        y = y - num / den;
    }
    y = y < 0.5 ? 0.5 : y;
    y = y > 2.0 ? 2.0 : y;
    return y;
}

// Squared distance between two axis-aligned boxes.
__device__ __forceinline__ double sphx_box_d2(double axl, double axh, double ayl, double ayh,
                                              double azl, double azh, double bxl, double bxh,
                                              double byl, double byh, double bzl, double bzh) {
    double dx = fmax(0.0, fmax(bxl - axh, axl - bxh));
    double dy = fmax(0.0, fmax(byl - ayh, ayl - byh));
    double dz = fmax(0.0, fmax(bzl - azh, azl - bzh));
    return dx * dx + dy * dy + dz * dz;
}

// Conservative reach test: true when a pair within `reach` may exist.
__device__ __forceinline__ bool sphx_within(double d2, double reach) {
    return d2 <= reach * reach * (1.0 + 1e-9);
}

// Compensated FP32 accumulation (Kahan): s + c carries the running sum.
struct KahanF {
    float s, c;
    __device__ __forceinline__ void init() { s = 0.f; c = 0.f; }
    __device__ __forceinline__ void add(float x) {
#pragma clang fp contract(off)
        float y = x - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

__device__ __forceinline__ double wave_max_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_min_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max_f(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_min_f(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// ---------------------------------------------------------------------------
// Device-side tables shared by the kernels (plain struct passed by value)
// ---------------------------------------------------------------------------
struct Soa {
    double *x, *y, *z, *vx, *vy, *vz, *u, *m, *h, *eps, *ut_prev;
    double *ax, *ay, *az, *gx, *gy, *gz, *rho, *P, *c, *divv, *curlv, *du, *pot;
    int32_t *k, *id, *run;
    int32_t *bin;  // block time-step level of each particle (block mode)
};

struct Tables {
    int32_t nb, ns, R;
    int32_t boff;                               // first bucket of a restricted launch
    const int32_t *b_beg, *b_cnt, *b_run;       // buckets
    const int32_t *s_bbeg, *s_bend;             // supers: bucket ranges
    const int32_t *r_sbeg, *r_send;             // run -> super range
    const int32_t *r_pbeg, *r_pend;             // run -> particle range
    const int32_t *r_lcut;                      // run -> coarsest Morton level a bucket may span
    double *bxl, *bxh, *byl, *byh, *bzl, *bzh, *bhmax, *bemax;  // bucket boxes
    double *sxl, *sxh, *syl, *syh, *szl, *szh, *shmax;          // super boxes
    const uint64_t *keys;                       // sorted Morton keys of the particles
    int32_t mbits;                              // Morton bits per axis in the keys
};

// Counters accumulated on device (int64 slots).
enum {
    CNT_H_ITERS_MAX = 0,
    CNT_H_ITERS_SUM,
    CNT_H_CAPPED,
    CNT_COOL_CAPPED,
    CNT_CAND,
    CNT_BH_OVERFLOW,
    CNT_NONFINITE,
    CNT_BIN_FLOOR,
    CNT_COOL_ITERS,   // bisection iterations of the implicit cooling solve
    CNT_COOL_BISECT,  // particles whose cooling solve bisected (not an early exit)
    CNT_NBR_OVF,      // buckets whose neighbour list overflowed (walked instead)
    CNT_SLOTS = 11
};

// Step phases timed with HIP events when Engine::timing is on.
enum {
    PH_SORT = 0,   // half kick + drift + Morton sort + gather
    PH_SMOOTH,     // boxes + smoothing-length iteration
    PH_DENS,       // boxes with h + density + EOS
    PH_GRAV,       // tile boxes + gravity
    PH_DIVCURL,
    PH_HYDRO,
    PH_FINISH,     // combine + half kick + u update + cooling
    PH_GRAVK,      // the direct-sum gravity kernel alone (inside PH_GRAV)
    PH_COUNT
};

// Per-run parameters on device.  dt is the step size of the next step; with
// `adaptive` set, the end of every step replaces it by the global time-step
// criterion on that step's forces (dt_parts keeps the four limits).  t is the
// simulated time, advanced by the dt each step used.  The device copy is the
// authority for dt and t once steps run (Engine::pull_run_params).
struct RunParams {
    double dt;
    int32_t use_cool;
    int32_t eps_from_h;
    int32_t adaptive;
    int32_t pad;
    double t;
    double c_cfl, dt_min, dt_max;
    double dt_parts[4];  // dt_f, dt_kin, dt_cfl, dt_dens of the last evaluation
};

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup w runs on
// XCD w % 8).  Spatially adjacent buckets share neighbour candidates, so give
// each XCD (and its L2) a contiguous range of buckets instead of every 8th one:
// XCD x handles [x q + min(x, r), ...) with q = nblk / 8, r = nblk % 8.  A
// bijection on [0, nblk); it changes which workgroup computes a bucket, never
// the bucket's arithmetic.
__device__ __forceinline__ int xcd_block(int bid, int nblk) {
    constexpr int X = 8;
    int q = nblk / X, r = nblk % X;
    int x = bid % X, k = bid / X;
    return x * q + (x < r ? x : r) + k;
}

// 2-D (u, rho) cooling table on the device
struct GridDev {
    int64_t nu = 0, nrho = 0;
    const double *u_nodes = nullptr, *rho_nodes = nullptr, *du = nullptr;
    double dt_ref_s = 0, unit_u_cgs = 0, unit_rho_cgs = 0, unit_time_s = 0;
};

// The analytic H/He network on the device: its parameters and the six
// background rates (passed by value).
struct NetDev {
    double XH = 0, tol = 0, unit_u_cgs = 0, unit_rho_cgs = 0, unit_time_s = 0;
    double rf[6] = {};
    int32_t max_iter = 0, max_bracket = 0, on = 0;
};

// Device view of a 5-D CHIMES table (sphx_chimes_table).  ion: per-particle
// species fractions indexed by the particle's caller id (plans only; null in
// batches and stateless calls use their own output).
struct ChimesDev {
    int32_t n_kpc = 0, n_nh = 0, n_z = 0, n_t = 0, n_m = 0, n_time = 0, z_index = 0;
    const float *kpc = nullptr, *nh = nullptr, *time_s = nullptr;
    const float *u_evol = nullptr, *mu = nullptr, *metalz = nullptr;
    double xh = 0, m_h = 0, u_max_cgs = 0, kpc_cm = 0;
    double unit_length_cm = 0, unit_rho_cgs = 0, unit_u_cgs = 0, unit_time_s = 0;
    float *ion = nullptr;
};

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
// Gravity launch configuration: particles per lane, j-tile size, j split.
struct GravConfig {
    int ipl = 2, tj = 512, js = 1;
};

// Gravity i-block / j-tile tables, rebuilt when runs or configuration change.
struct GravHost {
    int ipl = 0, tj = 0;
    int nblocks = 0, ntiles = 0;
    int32_t *dev = nullptr;  // ib_beg | ib_cnt | ib_run | r_tbeg
    std::vector<int32_t> run_ib0;  // first i-block of each run (+ total at the end)
    int64_t dev_cap = 0;
    double *tbox = nullptr;  // 8 doubles per tile
    int64_t tbox_cap = 0;
    double *part = nullptr;  // j-split partial sums
    int64_t part_cap = 0;
    std::vector<int64_t> sizes;
    struct SymHost *sym = nullptr;  // tables of the symmetric sum (configuration ipl = 8)
};
void sphx_sym_host_free(struct SymHost *h);
int64_t sphx_sym_run_bytes(int64_t n, int S);
int64_t sphx_sym_min_bytes(int64_t n, int S);
int sphx_set_cell_level(struct Engine &E, int32_t bias);
int sphx_pair_probe_launch(int64_t blocks, int iters, float *out, hipStream_t st);
int sphx_sym_reserve(struct Engine &E, int S, int64_t budget_bytes);
bool sphx_sym_fits(const struct Engine &E, const std::vector<int64_t> &sizes, int S);

// Morton bits per axis for an engine holding up to max_runs runs (sphx_sort.hip).
int sphx_morton_bits(int max_runs);

struct Engine {
    int device = 0;
    int mode = SPHX_MIXED;
    sphx_params p{};
    bool has_cool = false;
    sphx_cooling cool{};
    double *cool_T = nullptr, *cool_L = nullptr;
    GridDev grid{};            // optional 2-D cooling table (runs with use_cool == 2)
    double *grid_buf = nullptr;
    int64_t grid_cap = 0;
    int set_cool_grid(const sphx_cool_grid *g, hipStream_t st);
    NetDev net{};              // optional analytic H/He network (runs with use_cool == 4)
    ChimesDev chimes{};        // optional 5-D CHIMES table (runs with use_cool == 3)
    float *chimes_buf = nullptr;
    int64_t chimes_cap = 0;
    float *chimes_ion = nullptr;  // cap x n_m species fractions (plans)
    int64_t chimes_ion_cap = 0;
    int set_chimes(const sphx_chimes_table *t, bool keep_ions, hipStream_t st);

    int64_t cap = 0, n = 0;
    int R = 0;
    std::vector<int64_t> run_n;  // host copy of run sizes (packed order)
    std::vector<RunParams> run_params;

    Soa s{};
    // scratch
    uint64_t *keys = nullptr, *keys2 = nullptr;
    int32_t *perm = nullptr, *perm2 = nullptr;
    double *tmpd = nullptr;   // n doubles
    int32_t *tmpi = nullptr;  // n int32
    uint32_t *hist = nullptr; // radix histograms
    int64_t hist_cap = 0;
    double *bbox = nullptr;   // per-run bounding boxes (6 per run) + chunk partials
    int64_t bbox_cap = 0;
    int32_t *h_iters = nullptr;
    int64_t *counters = nullptr;  // CNT_SLOTS
    RunParams *d_run = nullptr;
    double *dtpart = nullptr;     // 5 extrema per bucket for the time-step criterion
    int64_t dtpart_cap = 0;

    // tables
    std::vector<int32_t> hb_beg, hb_cnt, hb_run, hs_bbeg, hs_bend, hr_sbeg, hr_send, hr_pbeg,
        hr_pend, hr_lcut;
    int32_t *tab_i32 = nullptr;  // packed device copy of the int tables
    int64_t tab_i32_cap = 0;
    double *tab_d = nullptr;     // box arrays
    int64_t tab_d_cap = 0;
    Tables t{};
    // neighbour lists of the density pass for the div/curl and hydro passes
    // (plans and batches; allocated with the capacity when want_lists is set)
    bool want_lists = false;
    bool use_lists = true;      // tuning knob: the walks give the same bits
    int32_t *nbr = nullptr;     // nbr_bcap buckets x SPHX_NBR_MAX x 64
    int32_t *nbr_ovf = nullptr; // two per bucket (one per half of a half-bucket walk)
    int64_t nbr_bcap = 0;
    PairRec *prec = nullptr;    // per particle of capacity
    float *fbal = nullptr;      // per particle of capacity
    bool lists_ready() const;
    int32_t *cand_off = nullptr, *cand_cnt = nullptr, *cand = nullptr;
    int64_t cand_cap = 0;
    int64_t cand_total = 0;
    int lcut_bias = 0;  // shift of the per-run bucket cut level (tuning only: sums do not change)

    // block time steps (plans): levels B (0: off), longest step, substep tick
    int blk_B = 0;
    double blk_dtmax = 0.0, blk_c_cfl = 0.25;
    int64_t blk_tick = 0;   // substeps since the last tick that every level shares
    int64_t blk_t1 = 0;     // > 0 while the forces of substep end t1 are computed
    bool blk_started = false;
    int32_t *bin_spare = nullptr;
    int block_open(hipStream_t st);
    int block_drift(double d, hipStream_t st);
    int block_close(int64_t t1, hipStream_t st);
    int block_substep(hipStream_t st);

    std::vector<void *> owned;  // every hipMalloc'd pointer

    GravConfig gcfg;
    GravHost grav;
    // Barnes-Hut gravity in the step (theta > 0; single-run mixed plans)
    struct BhHost *bh = nullptr;
    double bh_theta = 0.0;

    // Launch restriction to one run (bootstrap of a newly admitted run):
    // bucket range [rb0, rb1), particle range [rp0, rp1), run index rrun.
    int32_t rrun = -1, rb0 = 0, rb1 = 0, rp0 = 0, rp1 = 0;
    int restrict_to(int run);  // run < 0: whole set
    Tables tview() const {
        Tables v = t;
        v.boff = rrun >= 0 ? rb0 : 0;
        v.keys = keys;
        v.mbits = morton_bits();
        return v;
    }
    unsigned nb_launch() const { return (unsigned)(rrun >= 0 ? rb1 - rb0 : t.nb); }
    int max_runs = 1;          // run id bits are reserved for this many runs
    // bits per axis of the Morton keys: 21 for a lone run, fewer when run ids
    // share the key (the layout k_morton writes and the bucket cuts read)
    int morton_bits() const { return sphx_morton_bits(max_runs); }

    // Step graphs: one captured hipGraph of one step per pointer parity (the
    // sort swaps state and spare arrays, so consecutive steps alternate
    // between two pointer assignments).  Valid for the current run set.
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    bool graph_valid[2] = {false, false};
    int gpar = 0;
    bool use_graph = false;
    void drop_graph();
    void swap_sorted();
    int step_body(hipStream_t st);
    int step_graph(hipStream_t st);
    // n graph steps on a private non-blocking stream, ordered after and
    // before the caller's stream with events (capture needs a real stream)
    hipStream_t gstream = nullptr;
    hipEvent_t gev_in = nullptr, gev_out = nullptr;
    int step_graphs(hipStream_t caller, int n);

    // optional per-phase HIP-event timing of the step
    bool timing = false;
    hipEvent_t tev[2 * 8] = {};
    bool tmarked[8] = {};      // phase recorded since the last collection
    double phase_ms[8] = {};
    int64_t phase_n[8] = {};
    int tmark(hipStream_t st, int phase, int end);
    int tcollect();

    int steps = 0;
    bool forces_valid = false;

    ~Engine();
    int alloc(void **ptr, size_t bytes);
    int reserve(int64_t capacity);
    // Rebuild the run tables; the upload is ordered after the work already
    // queued on st (which may still read the old tables), then waited for.
    int set_runs(const std::vector<int64_t> &sizes, hipStream_t st = nullptr);
    int upload_run_params(hipStream_t st);
    int pull_run_params(hipStream_t st);  // device dt, t, dt_parts -> host copy

    // pipeline pieces (mixed mode)
    int sort_and_permute(hipStream_t st, bool gather_state);
    int build_boxes(hipStream_t st, bool with_h);
    int smoothing(hipStream_t st, bool predictor);
    int build_candidates(hipStream_t st);
    int density_eos(hipStream_t st, bool lists = false);
    // groups per bucket of the density walk (1 or 2)
    int walk_halves() const;
    int gravity(hipStream_t st, bool with_pot);
    int divcurl(hipStream_t st, bool lists = false);
    int hydro(hipStream_t st, bool lists = false);
    // div/curl and hydro in one pass over the lists (lists on, no Balsara)
    int divcurl_hydro_lists(hipStream_t st);
    int forces(hipStream_t st, bool predictor, bool with_pot);

    // plain FP64 baseline pieces
    int plain_forces(hipStream_t st, bool predictor, bool with_pot);

    // integrator
    int kick_drift(hipStream_t st);
    int finish_step(hipStream_t st);
    int bootstrap_finish(hipStream_t st);
    int global_dt(hipStream_t st);  // t += dt; adaptive runs: dt <- criterion
    int step(hipStream_t st, int nsteps);

    int read_counters(sphx_counters *out);
    int diagnostics(hipStream_t st, bool with_pe, sphx_diag *out);
};

// kNN smoothing lengths into E.s.h (sorted order), from the search radius 2 h
// already in E.s.h (sphx_hydro.hip)
int sphx_knn_pass(Engine &E, int kth, hipStream_t st);

// Exact checkpoint of one resident run (sphx_io.hip)
#define SPHX_CKPT_MAGIC "SPHXRUN1"
#define SPHX_CKPT_VERSION 2  // 2: sphx_params carries h_min
#define SPHX_CKPT_ND 24  // float64 arrays: every double array of Soa
#define SPHX_CKPT_NI 2   // int32 arrays: k, id
struct CkptHeader {
    char magic[8];
    int32_t version, mode;
    int64_t n, steps;
    sphx_params p;
    RunParams rp;
    GravConfig g;
    int32_t seed_pending;  // the next plan step's dt seeds the adaptive step
    double bh_theta;
    int32_t morton_bits, reserved[7];
};
int sphx_ckpt_write(const char *path, const CkptHeader &hdr, Engine &E, int64_t p0,
                    hipStream_t st);
int sphx_ckpt_read_header(const char *path, CkptHeader *hdr);
int sphx_ckpt_upload(const char *path, Engine &E, int64_t p0, int64_t n, hipStream_t st);

// Barnes-Hut tree (sphx_bh.hip)
struct BhHost;
struct BhStats {
    int64_t nodes, depth, interactions, overflow, max_fill, groups;
    double build_ms, walk_ms;  // device time (HIP events) of the build and the walk
};
BhHost *sphx_bh_host_new();
void sphx_bh_host_free(BhHost *h);
int sphx_bh_run(Engine &E, BhHost &H, double theta, bool with_pot, hipStream_t st, BhStats *out);

// launchers implemented in the .hip files
int sphx_gravity_run(Engine &E, const GravConfig &cfg, bool with_pot, hipStream_t st);
int sphx_gravity_prepare(Engine &E, const GravConfig &cfg, hipStream_t st);
int sphx_launch_morton(Engine &E, hipStream_t st);
int sphx_radix_sort(uint64_t *keys, int32_t *vals, uint64_t *keys_tmp, int32_t *vals_tmp,
                    uint32_t *hist, int64_t n, int key_bits, hipStream_t st);
// Time-step criterion over n particles (acc and acc_g with element strides
// as/gs); out[0..4] = dt, dt_f, dt_kin, dt_cfl, dt_dens.
int sphx_global_dt_launch(const double *h, const double *c, const double *acc, int64_t as,
                          const double *acc_g, int64_t gs, const double *divv, int64_t n,
                          double c_cfl, double dt_min, double dt_max, double *scratch,
                          int64_t scratch_len, double *out, hipStream_t st);
int sphx_diag_reduce(const double *x, const double *y, const double *z, const double *vx,
                     const double *vy, const double *vz, const double *m, const double *u,
                     const double *pot, int64_t n, double *scratch, sphx_diag *out,
                     hipStream_t st);

// plain FP64 kernels (SoA inputs, rows [r0, r1))
int plain_smoothing(const double *x, const double *y, const double *z, const double *h_prev,
                    const int32_t *k_prev, const double *divv, const double *dt_p, double dt,
                    const int32_t *run, const sphx_params &p, int64_t n, int64_t r0, int64_t r1,
                    double *h_out, int32_t *k_out, int32_t *it_out, int64_t *counters,
                    hipStream_t st);
int plain_density(const double *x, const double *y, const double *z, const double *m,
                  const double *h, int64_t n, int64_t r0, int64_t r1, double *rho,
                  int32_t *nngb, hipStream_t st);
int plain_eos(const double *rho, const double *u, double gamma, int64_t n, double *P,
              double *c, hipStream_t st);
int plain_divcurl(const double *x, const double *y, const double *z, const double *vx,
                  const double *vy, const double *vz, const double *rho, const double *m,
                  const double *h, int64_t n, int64_t r0, int64_t r1, double *divv,
                  double *curlv, hipStream_t st);
int plain_acc_sph(const Soa &s, const sphx_params &p, int64_t n, int64_t r0, int64_t r1,
                  double *ax, double *ay, double *az, hipStream_t st);
int plain_du(const Soa &s, const sphx_params &p, int64_t n, int64_t r0, int64_t r1, double *du,
             hipStream_t st);
int plain_gravity(const double *x, const double *y, const double *z, const double *m,
                  const double *eps, double G, int64_t n, int64_t r0, int64_t r1, double *ax,
                  double *ay, double *az, double *pot, hipStream_t st);
int plain_neighbors(const double *x, const double *y, const double *z, const double *h,
                    int64_t n, int64_t r0, int64_t r1, int64_t *offsets, int32_t *idx,
                    hipStream_t st);

// cooling (FP64 implicit solve, shared by every mode)
struct CoolDev {
    double unit_u_cgs, unit_rho_cgs, unit_time_s, XH, mu, gamma, heat, tol;
    int32_t max_iter, max_bracket, ntab;
    const double *T, *L;
};
int sphx_network_dev(const sphx_network *net, NetDev *out);
int sphx_net_launch(const NetDev &nd, const double *rho, const double *u_in, double dt,
                    int64_t n, double *u_out, int32_t *iters, int64_t *counters, hipStream_t st);
int sphx_grid_apply_launch(const GridDev &g, const double *rho, const double *u_in, double dt,
                           int64_t n, double *u_out, hipStream_t st);
int sphx_grid_build_launch(const GridDev &g, const CoolDev &cd, double *du, hipStream_t st);
int sphx_chimes_apply_launch(const ChimesDev &c, const double *pos, const double *rho,
                             const double *u_in, double dt, int64_t n, double *u_out,
                             float *ion_out, double *mu_out, hipStream_t st);
int sphx_check_chimes(const sphx_chimes_table *t, hipStream_t st);
ChimesDev sphx_chimes_view(const sphx_chimes_table *t);
int sphx_cool_launch(const CoolDev &cd, const double *rho, const double *u_in, double dt,
                     const double *dt_p, const int32_t *run, const RunParams *rp, int64_t n,
                     double *u_out, int32_t *iters, int64_t *counters, hipStream_t st);

// AoS <-> SoA helpers
int sphx_aos_to_soa(const double *aos, int64_t n, double *x, double *y, double *z,
                    hipStream_t st);
int sphx_soa_to_aos(const double *x, const double *y, const double *z, int64_t n, double *aos,
                    hipStream_t st);
