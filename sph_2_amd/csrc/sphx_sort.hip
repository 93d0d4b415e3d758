// Morton keys, stable LSD radix sort, permutation gather and bucket boxes.
//
// Replaces the reference's neighbour structures (counting-sort cells of
// Particles_in_BH_Tree_II/Evrard_test/ngb_v3.h:58-117, atomics-filled grid
// cells of Particles_in_Gridcells/final_revision_uniform_grid/
// hfvCLibs_v7_grid_v0.h:680-708) with Morton-ordered 64-particle buckets:
// one wavefront per bucket, bucket and super-bucket AABBs for culling.
#include <utility>

#include "sphx_internal.h"

namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_TILE = SPHX_RS_TILE;
constexpr int RS_ITEMS = RS_TILE / RS_THREADS;
static_assert(RS_ITEMS * RS_THREADS == RS_TILE, "a tile is whole rounds of the workgroup");

__device__ __forceinline__ uint64_t spread3(uint64_t v) {
    // Spread the low 21 bits of v so bit b lands at bit 3b.
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

// Bounding box of each super-bucket straight from the particles (before the
// sort, in the current order; supers never straddle runs).
__global__ __launch_bounds__(256) void k_super_box_pos(Tables t, const double *__restrict__ x,
                                                       const double *__restrict__ y,
                                                       const double *__restrict__ z,
                                                       double *__restrict__ out /* 6 per super */) {
    __shared__ double part[4][6];
    const int sidx = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int b0 = t.s_bbeg[sidx], b1 = t.s_bend[sidx];
    const int p0 = t.b_beg[b0];
    const int p1 = t.b_beg[b1 - 1] + t.b_cnt[b1 - 1];
    double v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (int i = p0 + tid; i < p1; i += 256) {
        const double a = x[i], b = y[i], c = z[i];
        v[0] = fmin(v[0], a); v[1] = fmax(v[1], a);
        v[2] = fmin(v[2], b); v[3] = fmax(v[3], b);
        v[4] = fmin(v[4], c); v[5] = fmax(v[5], c);
    }
    for (int k = 0; k < 6; k += 2) {
        v[k] = wave_min_d(v[k]);
        v[k + 1] = wave_max_d(v[k + 1]);
    }
    if (lane == 0)
        for (int k = 0; k < 6; ++k) part[w][k] = v[k];
    __syncthreads();
    if (tid < 6) {
        double r = part[0][tid];
        for (int q = 1; q < 4; ++q) r = (tid & 1) ? fmax(r, part[q][tid]) : fmin(r, part[q][tid]);
        out[6 * sidx + tid] = r;
    }
}

// Per-run bounding box from the super boxes; one wavefront per run.
__global__ void k_run_box(Tables t, const double *__restrict__ sbox, double *__restrict__ rbox) {
    int r = blockIdx.x;
    int lane = threadIdx.x;
    double v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (int s = t.r_sbeg[r] + lane; s < t.r_send[r]; s += 64) {
        for (int a = 0; a < 6; a += 2) {
            v[a] = fmin(v[a], sbox[6 * s + a]);
            v[a + 1] = fmax(v[a + 1], sbox[6 * s + a + 1]);
        }
    }
    for (int a = 0; a < 6; a += 2) {
        v[a] = wave_min_d(v[a]);
        v[a + 1] = wave_max_d(v[a + 1]);
    }
    if (lane == 0)
        for (int a = 0; a < 6; ++a) rbox[6 * r + a] = v[a];
}

// key = run << (3*bits) | interleave(ix, iy, iz); ix = floor((x - lo) * scale)
// clamped to [0, 2^bits - 1], scale = 2^bits / (extent * (1 + 2^-30)) over the
// run's cubic box.  All FP64, separately rounded.
__global__ void k_morton(const double *__restrict__ x, const double *__restrict__ y,
                         const double *__restrict__ z, const int32_t *__restrict__ run,
                         const double *__restrict__ rbox, int64_t n, int bits,
                         uint64_t *__restrict__ keys, int32_t *__restrict__ vals) {
#pragma clang fp contract(off)
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int r = run ? run[i] : 0;
    const double *b = rbox + 6 * r;
    double ex = b[1] - b[0], ey = b[3] - b[2], ez = b[5] - b[4];
    double ext = fmax(ex, fmax(ey, ez));
    double cells = (double)(1ull << bits);
    double scale = ext > 0.0 ? cells / (ext * (1.0 + 0x1p-30)) : 0.0;
    uint64_t maxc = (1ull << bits) - 1ull;
    double fx = (x[i] - b[0]) * scale, fy = (y[i] - b[2]) * scale, fz = (z[i] - b[4]) * scale;
    uint64_t ix = fx <= 0.0 ? 0ull : (uint64_t)fx;
    uint64_t iy = fy <= 0.0 ? 0ull : (uint64_t)fy;
    uint64_t iz = fz <= 0.0 ? 0ull : (uint64_t)fz;
    ix = ix > maxc ? maxc : ix;
    iy = iy > maxc ? maxc : iy;
    iz = iz > maxc ? maxc : iz;
    uint64_t key = (spread3(ix) << 2) | (spread3(iy) << 1) | spread3(iz);
    key |= ((uint64_t)r) << (3 * bits);
    keys[i] = key;
    vals[i] = (int32_t)i;
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t *__restrict__ keys,
                                                        int64_t n, int shift,
                                                        uint32_t *__restrict__ hist,
                                                        int nblocks) {
    __shared__ uint32_t h[256];
    int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; ++r) {
        int64_t i = base + r * RS_THREADS + tid;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)tid * nblocks + blockIdx.x] = h[tid];
}

// The scatter offsets are the exclusive scan of the block histograms in
// digit-major order (entry d * nblocks + b counts digit d in block b).  Two
// kernels with one workgroup per digit: the first sums each digit's row, the
// second scans its row on top of the totals of the digits before it.
__global__ __launch_bounds__(256) void k_rs_digit_total(const uint32_t *__restrict__ hist,
                                                        int nblocks, uint32_t *__restrict__ tot) {
    __shared__ uint32_t red[256];
    const int d = blockIdx.x, tid = threadIdx.x;
    uint32_t s = 0;
    for (int b = tid; b < nblocks; b += 256) s += hist[(int64_t)d * nblocks + b];
    red[tid] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) tot[d] = red[0];
}

__global__ __launch_bounds__(256) void k_rs_scan_rows(uint32_t *__restrict__ hist, int nblocks,
                                                      const uint32_t *__restrict__ tot) {
    __shared__ uint32_t part[256];
    const int d = blockIdx.x, tid = threadIdx.x;
    part[tid] = tid < d ? tot[tid] : 0u;  // the digits before this one
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) part[tid] += part[tid + o];
        __syncthreads();
    }
    const uint32_t base = part[0];
    __syncthreads();
    // each thread a consecutive piece of the row
    const int per = (nblocks + 255) / 256;
    const int b0 = min(nblocks, tid * per), b1 = min(nblocks, b0 + per);
    uint32_t *row = hist + (int64_t)d * nblocks;
    uint32_t s = 0;
    for (int b = b0; b < b1; ++b) s += row[b];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t run = base + (tid > 0 ? part[tid - 1] : 0u);
    for (int b = b0; b < b1; ++b) {
        uint32_t v = row[b];
        row[b] = run;
        run += v;
    }
}

// Stable scatter: items keep their order inside each digit.  Items of a tile
// are taken in rounds of 256 (thread order); inside a wavefront, peers with
// the same digit are found by matching the 8 digit bits with ballots.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(
    const uint64_t *__restrict__ kin, const int32_t *__restrict__ vin, uint64_t *__restrict__ kout,
    int32_t *__restrict__ vout, int64_t n, int shift, const uint32_t *__restrict__ hist,
    int nblocks) {
    __shared__ uint32_t base[256];
    __shared__ uint32_t runc[256];
    __shared__ uint32_t wcnt[RS_THREADS / 64][256];
    int tid = threadIdx.x;
    int lane = tid & 63;
    int w = tid >> 6;
    base[tid] = hist[(int64_t)tid * nblocks + blockIdx.x];
    runc[tid] = 0;
    uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int64_t tile = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; ++r) {
        for (int q = 0; q < RS_THREADS / 64; ++q) wcnt[q][tid] = 0;
        __syncthreads();
        int64_t i = tile + r * RS_THREADS + tid;
        bool valid = i < n;
        uint64_t key = valid ? kin[i] : 0ull;
        int32_t val = valid ? vin[i] : 0;
        uint32_t d = (uint32_t)((key >> shift) & 255u);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t m = __ballot(valid && ((d >> b) & 1u));
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = base[d] + runc[d] + rank;
            for (int q = 0; q < w; ++q) off += wcnt[q][d];
            kout[off] = key;
            vout[off] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int q = 0; q < RS_THREADS / 64; ++q) add += wcnt[q][tid];
        runc[tid] += add;
        // next round's clearing of wcnt follows a barrier at its top
        __syncthreads();
    }
}

struct GatherList {
    const double *src[16];
    double *dst[16];
    const int32_t *isrc[4];
    int32_t *idst[4];
    int nd, ni;
};

__global__ void k_gather(GatherList g, const int32_t *__restrict__ perm, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t p = perm[i];
    for (int a = 0; a < g.nd; ++a) g.dst[a][i] = g.src[a][p];
    for (int a = 0; a < g.ni; ++a) g.idst[a][i] = g.isrc[a][p];
}

// Box and largest h and softening of each 64-particle bucket: one wave each.
__global__ __launch_bounds__(64) void k_bucket_boxes(Tables t, const double *__restrict__ x,
                                                     const double *__restrict__ y,
                                                     const double *__restrict__ z,
                                                     const double *__restrict__ h,
                                                     const double *__restrict__ eps, int with_h) {
    const int b = blockIdx.x, lane = threadIdx.x;
    const int cnt = t.b_cnt[b];
    const int i = t.b_beg[b] + lane;
    const bool ok = lane < cnt;
    double a = ok ? x[i] : INFINITY, bb = ok ? y[i] : INFINITY, c = ok ? z[i] : INFINITY;
    const double xl = wave_min_d(a), yl = wave_min_d(bb), zl = wave_min_d(c);
    a = ok ? x[i] : -INFINITY; bb = ok ? y[i] : -INFINITY; c = ok ? z[i] : -INFINITY;
    const double xh = wave_max_d(a), yh = wave_max_d(bb), zh = wave_max_d(c);
    double hm = 0.0, em = 0.0;
    if (with_h) {
        hm = wave_max_d(ok ? h[i] : 0.0);
        em = wave_max_d(ok ? (eps ? eps[i] : h[i]) : 0.0);
    }
    if (lane == 0) {
        t.bxl[b] = xl; t.bxh[b] = xh; t.byl[b] = yl; t.byh[b] = yh;
        t.bzl[b] = zl; t.bzh[b] = zh;
        if (with_h) { t.bhmax[b] = hm; t.bemax[b] = em; }
    }
}

// Box and largest h of each super (up to SPHX_SUPER consecutive buckets) from
// its buckets' boxes: lane k takes buckets k, k + 64, ... of the super.
__global__ __launch_bounds__(64) void k_super_boxes(Tables t, int with_h) {
    const int sidx = blockIdx.x, lane = threadIdx.x;
    double S[7] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY, 0.0};
    for (int b = t.s_bbeg[sidx] + lane; b < t.s_bend[sidx]; b += 64) {
        S[0] = fmin(S[0], t.bxl[b]); S[1] = fmax(S[1], t.bxh[b]);
        S[2] = fmin(S[2], t.byl[b]); S[3] = fmax(S[3], t.byh[b]);
        S[4] = fmin(S[4], t.bzl[b]); S[5] = fmax(S[5], t.bzh[b]);
        if (with_h) S[6] = fmax(S[6], t.bhmax[b]);
    }
    const double xl = wave_min_d(S[0]), xh = wave_max_d(S[1]), yl = wave_min_d(S[2]),
                 yh = wave_max_d(S[3]), zl = wave_min_d(S[4]), zh = wave_max_d(S[5]);
    const double hm = with_h ? wave_max_d(S[6]) : 0.0;
    if (lane == 0) {
        t.sxl[sidx] = xl; t.sxh[sidx] = xh; t.syl[sidx] = yl; t.syh[sidx] = yh;
        t.szl[sidx] = zl; t.szh[sidx] = zh;
        if (with_h) t.shmax[sidx] = hm;
    }
}

__global__ void k_aos_to_soa(const double *__restrict__ a, int64_t n, double *__restrict__ x,
                             double *__restrict__ y, double *__restrict__ z) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = a[3 * i];
    y[i] = a[3 * i + 1];
    z[i] = a[3 * i + 2];
}

__global__ void k_soa_to_aos(const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, int64_t n, double *__restrict__ a) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    a[3 * i] = x[i];
    a[3 * i + 1] = y[i];
    a[3 * i + 2] = z[i];
}

inline unsigned grid_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace

int sphx_aos_to_soa(const double *aos, int64_t n, double *x, double *y, double *z,
                    hipStream_t st) {
    if (n <= 0) return SPHX_OK;
    k_aos_to_soa<<<grid_for(n, 256), 256, 0, st>>>(aos, n, x, y, z);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

int sphx_soa_to_aos(const double *x, const double *y, const double *z, int64_t n, double *aos,
                    hipStream_t st) {
    if (n <= 0) return SPHX_OK;
    k_soa_to_aos<<<grid_for(n, 256), 256, 0, st>>>(x, y, z, n, aos);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

// hist holds 256 x (nblocks + 1) entries: the block histograms, then the
// digit totals.
int sphx_radix_sort(uint64_t *keys, int32_t *vals, uint64_t *keys_tmp, int32_t *vals_tmp,
                    uint32_t *hist, int64_t n, int key_bits, hipStream_t st) {
    if (n <= 1) return SPHX_OK;
    int nblocks = (int)((n + RS_TILE - 1) / RS_TILE);
    uint32_t *tot = hist + (int64_t)256 * nblocks;
    int passes = (key_bits + 7) / 8;
    uint64_t *ka = keys, *kb = keys_tmp;
    int32_t *va = vals, *vb = vals_tmp;
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        k_rs_hist<<<nblocks, RS_THREADS, 0, st>>>(ka, n, shift, hist, nblocks);
        k_rs_digit_total<<<256, 256, 0, st>>>(hist, nblocks, tot);
        k_rs_scan_rows<<<256, 256, 0, st>>>(hist, nblocks, tot);
        k_rs_scatter<<<nblocks, RS_THREADS, 0, st>>>(ka, va, kb, vb, n, shift, hist, nblocks);
        SPHX_HIP(hipGetLastError());
        uint64_t *tk = ka; ka = kb; kb = tk;
        int32_t *tv = va; va = vb; vb = tv;
    }
    if (ka != keys) {
        SPHX_HIP(hipMemcpyAsync(keys, ka, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, st));
        SPHX_HIP(hipMemcpyAsync(vals, va, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
    }
    return SPHX_OK;
}

static int run_bits(int max_runs) {
    int rb = 0;
    while ((1 << rb) < max_runs) ++rb;
    return rb;
}

// Bits per axis: 21 for a lone run; fewer when run ids share the 63-bit key.
// Depends on the engine's run capacity only, so a run sorts the same way
// whether it is alone or packed with others.
int sphx_morton_bits(int max_runs) {
    int b = (63 - run_bits(max_runs)) / 3;
    return b > 21 ? 21 : b;
}

static int run_key_bits(int max_runs) { return 3 * sphx_morton_bits(max_runs) + run_bits(max_runs); }

int sphx_launch_morton(Engine &E, hipStream_t st) {
    Tables &t = E.t;
    double *sbox = E.bbox;                 // 6 per super
    double *rbox = E.bbox + 6 * (int64_t)t.ns;  // 6 per run
    k_super_box_pos<<<t.ns, 256, 0, st>>>(t, E.s.x, E.s.y, E.s.z, sbox);
    k_run_box<<<t.R, 64, 0, st>>>(t, sbox, rbox);
    k_morton<<<grid_for(E.n, 256), 256, 0, st>>>(E.s.x, E.s.y, E.s.z,
                                                E.max_runs > 1 ? E.s.run : nullptr, rbox, E.n,
                                                sphx_morton_bits(E.max_runs), E.keys, E.perm);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

// Persistent arrays gathered by the sort, and the derived arrays that serve
// as their destinations (recomputed after every sort).  After the gather the
// roles swap; applying the swap twice restores the original assignment.
static void sort_roles(Engine &E, double **dst[12], double **spare[12], int32_t **idst[4],
                       int32_t **ispare[4]) {
    Soa &s = E.s;
    double **d[12] = {&s.x, &s.y, &s.z, &s.vx, &s.vy, &s.vz, &s.u, &s.m,
                      &s.h, &s.eps, &s.ut_prev, &s.divv};
    double **sp[12] = {&s.gx, &s.gy, &s.gz, &s.rho, &s.P, &s.c, &s.curlv, &s.du,
                       &s.pot, &s.ax, &s.ay, &s.az};
    int32_t **id[4] = {&s.k, &s.id, &s.run, &s.bin};
    int32_t **isp[4] = {&E.tmpi, &E.perm2, &E.h_iters, &E.bin_spare};
    for (int a = 0; a < 12; ++a) { dst[a] = d[a]; spare[a] = sp[a]; }
    for (int a = 0; a < 4; ++a) { idst[a] = id[a]; ispare[a] = isp[a]; }
}

void Engine::swap_sorted() {
    double **d[12], **sp[12];
    int32_t **id[4], **isp[4];
    sort_roles(*this, d, sp, id, isp);
    for (int a = 0; a < 12; ++a) std::swap(*d[a], *sp[a]);
    for (int a = 0; a < 4; ++a) std::swap(*id[a], *isp[a]);
    gpar ^= 1;
}

int Engine::sort_and_permute(hipStream_t st, bool gather_state) {
    if (n == 0) return SPHX_OK;
    SPHX_CHECK(sphx_launch_morton(*this, st));
    SPHX_CHECK(sphx_radix_sort(keys, perm, keys2, perm2, hist, n, run_key_bits(max_runs), st));
    if (!gather_state) return SPHX_OK;
    double **d[12], **sp[12];
    int32_t **id[4], **isp[4];
    sort_roles(*this, d, sp, id, isp);
    GatherList g{};
    g.nd = 12;
    g.ni = 4;
    for (int a = 0; a < 12; ++a) {
        g.src[a] = *d[a];
        g.dst[a] = *sp[a];
    }
    for (int a = 0; a < 4; ++a) {
        g.isrc[a] = *id[a];
        g.idst[a] = *isp[a];
    }
    k_gather<<<grid_for(n, 256), 256, 0, st>>>(g, perm, n);
    SPHX_HIP(hipGetLastError());
    swap_sorted();
    return SPHX_OK;
}

int Engine::build_boxes(hipStream_t st, bool with_h) {
    if (n == 0) return SPHX_OK;
    const double *e = p.eps_from_h ? nullptr : s.eps;
    k_bucket_boxes<<<t.nb, 64, 0, st>>>(t, s.x, s.y, s.z, s.h, e, with_h ? 1 : 0);
    k_super_boxes<<<t.ns, 64, 0, st>>>(t, with_h ? 1 : 0);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}
