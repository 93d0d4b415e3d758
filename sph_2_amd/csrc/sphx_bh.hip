// Barnes-Hut gravity: Morton-built octree, monopole moments, group walk.
//
// Replaces the reference's iterative octree (Particles_in_BH_Tree_II/
// Evrard_test/bh_tree_iteration_v2.h: single-block atomic insertion build
// 149-367, per-warp stack walk 454-588, theta = 0.5) with a build that
// follows the sorted 63-bit Morton keys: every octree cell is a contiguous
// key range, and a cell's eight children are found by binary search on the
// next three key bits, one level per pass.  Moments (mass, centre of mass,
// max eps) are summed bottom-up in FP64.  The walk is per group (one
// wavefront): the groups are the largest cells holding at most 64 particles
// (unsplittable leaves of coincident particles are cut into runs of 64), so a
// group's bounding box is always inside one cell's cube -- unlike 64
// consecutive Morton indices, which can straddle a jump of the curve and then
// open most of the tree.  A cell is accepted for the whole group when its
// distance from the group's bounding box exceeds s/theta + |com - centre|
// (cell side s); accepted cells and the particles of opened leaves go to an
// LDS interaction list that is flushed through the FP32 spline-softened pair
// formula with Kahan-compensated sums.
#include <algorithm>

#include "sphx_internal.h"

namespace {

constexpr int BH_LEAF = 8;        // particles per leaf at most
// Cell stack and interaction list capacities per group (one wave).  A deep
// walk (small opening angle) opens many cells per step, and the larger stack
// keeps its batches of popped cells large; at larger angles the smaller
// stack's occupancy wins.  The walk is bound by how many groups a CU holds
// (LDS) and a SIMD holds (registers): a list of 128 entries and four entries
// evaluated together give four to five waves per SIMD, where a list of 512
// and batches of eight gave two to three (4M-particle Plummer sphere, theta
// 0.3 / 0.5 / 0.7: 45 / 13.8 / 7.5 ms against 60 / 18.1 / 10.1 ms; theta 0.3
// took 100 ms with the smaller stack).
constexpr int BH_SCAP = 1024;
constexpr int BH_SCAP_DEEP = 2048;
constexpr double BH_DEEP_THETA = 0.35;  // the larger stack below this angle
constexpr int BH_SLACK = 160;     // > 7 x 21 levels of pure depth-first growth
constexpr int BH_ICAP = 128;
constexpr int BH_BATCH = 4;       // list entries evaluated together
// leaf particles a lane lists per round: one round fills the list at most
constexpr int BH_ROUND = BH_ICAP / 64;
static_assert(BH_ICAP >= 64 && BH_ICAP % 64 == 0,
              "a batch of accepted cells (at most 64) must fit the list");
static_assert(BH_ROUND < 8, "leaf particles per lane and round are scanned in three bits");
static_assert(BH_SCAP > BH_SLACK + 7 * 64, "a full batch must fit above the slack");
static_assert(BH_SCAP_DEEP > BH_SCAP, "the deep-walk stack is the larger one");
constexpr int BHT = 256;

// What the walk reads of a cell, packed so that one cell costs two 32-byte
// loads: the monopole (FP64 centre of mass) and the opening/link record.
struct __align__(32) BhMono {
    double cx, cy, cz;
    float mass, emax;
};
struct __align__(32) BhLink {
    float size, delta;
    int32_t child, nchild, beg, cnt, pad0, pad1;
};

struct BhTree {
    int32_t *beg, *cnt, *child, *nchild, *level;
    double *mass, *cx, *cy, *cz, *gx, *gy, *gz, *size, *delta, *emax;
    BhMono *mono;
    BhLink *link;
};

__device__ __forceinline__ int digit_at(uint64_t key, int shift) { return (int)((key >> shift) & 7ull); }

// For each cell of the current level: the nine child boundaries (first index
// of each child digit) and the number of non-empty children.
__global__ void k_bh_split(BhTree T, int32_t f0, int32_t fc, const uint64_t *__restrict__ keys,
                           int level, int32_t *__restrict__ nk, int32_t *__restrict__ bounds) {
    int t = blockIdx.x * BHT + threadIdx.x;
    if (t >= fc) return;
    int i = f0 + t;
    int b = T.beg[i], c = T.cnt[i];
    int32_t *bd = bounds + 9 * t;
    if (c <= BH_LEAF || level >= 21) {
        nk[t] = 0;
        bd[0] = -1;
        return;
    }
    int shift = 3 * (20 - level);
    bd[0] = b;
    bd[8] = b + c;
    for (int d = 1; d < 8; ++d) {
        int lo = b, hi = b + c;  // first index with digit >= d
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (digit_at(keys[mid], shift) < d) lo = mid + 1;
            else hi = mid;
        }
        bd[d] = lo;
    }
    int k = 0;
    for (int d = 0; d < 8; ++d) k += bd[d + 1] > bd[d] ? 1 : 0;
    nk[t] = k;
}

// Exclusive scan of fc counts (one workgroup), total written to *total.
// Exclusive scan of the fc child counts of a level, in three launches of
// 1024-entry tiles: the tile sums, a scan of the tile sums (writing the
// total), then each tile's scan on top of its offset.
constexpr int BH_SCAN_TILE = 1024;

__global__ __launch_bounds__(256) void k_bh_scan_tiles(const int32_t *__restrict__ in, int32_t fc,
                                                       int32_t *__restrict__ tsum) {
    __shared__ int32_t red[256];
    const int t = blockIdx.x, tid = threadIdx.x;
    const int64_t b = (int64_t)t * BH_SCAN_TILE;
    int s = 0;
    for (int k = tid; k < BH_SCAN_TILE; k += 256)
        if (b + k < fc) s += in[b + k];
    red[tid] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) tsum[t] = red[0];
}

__global__ __launch_bounds__(1024) void k_bh_scan_sums(int32_t *__restrict__ tsum, int ntiles,
                                                       int32_t *__restrict__ total) {
    __shared__ int32_t part[1024];
    const int tid = threadIdx.x;
    const int per = (ntiles + 1023) / 1024;
    const int b = min(ntiles, tid * per), e = min(ntiles, b + per);
    int s = 0;
    for (int i = b; i < e; ++i) s += tsum[i];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = tid > 0 ? part[tid - 1] : 0;
    for (int i = b; i < e; ++i) {
        int v = tsum[i];
        tsum[i] = run;
        run += v;
    }
    if (tid == 1023) *total = part[1023];
}

__global__ __launch_bounds__(256) void k_bh_scan_apply(const int32_t *__restrict__ in, int32_t fc,
                                                       const int32_t *__restrict__ tsum,
                                                       int32_t *__restrict__ out) {
    __shared__ int32_t part[256];
    const int t = blockIdx.x, tid = threadIdx.x;
    const int64_t b = (int64_t)t * BH_SCAN_TILE + 4 * tid;  // four consecutive entries per thread
    int v[4], s = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = b + k < fc ? in[b + k] : 0;
        s += v[k];
    }
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int u = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += u;
        __syncthreads();
    }
    int run = tsum[t] + (tid > 0 ? part[tid - 1] : 0);
    for (int k = 0; k < 4; ++k) {
        if (b + k < fc) out[b + k] = run;
        run += v[k];
    }
}

__global__ void k_bh_emit(BhTree T, int32_t f0, int32_t fc, const int32_t *__restrict__ off,
                          const int32_t *__restrict__ bounds, int32_t base, int level) {
    int t = blockIdx.x * BHT + threadIdx.x;
    if (t >= fc) return;
    int i = f0 + t;
    const int32_t *bd = bounds + 9 * t;
    if (bd[0] < 0) {
        T.child[i] = -1;
        T.nchild[i] = 0;
        return;
    }
    int o = base + off[t];
    int k = 0;
    for (int d = 0; d < 8; ++d) {
        if (bd[d + 1] > bd[d]) {
            T.beg[o + k] = bd[d];
            T.cnt[o + k] = bd[d + 1] - bd[d];
            T.level[o + k] = level + 1;
            ++k;
        }
    }
    T.child[i] = o;
    T.nchild[i] = k;
}

__device__ __forceinline__ uint32_t compact3(uint64_t v) {
    // gather every third bit of v (bit 3b -> bit b)
    v &= 0x1249249249249249ull;
    v = (v | (v >> 2)) & 0x10c30c30c30c30c3ull;
    v = (v | (v >> 4)) & 0x100f00f00f00f00full;
    v = (v | (v >> 8)) & 0x1f0000ff0000ffull;
    v = (v | (v >> 16)) & 0x1f00000000ffffull;
    v = (v | (v >> 32)) & 0x1fffffull;
    return (uint32_t)v;
}

// Moments of the cells of one level (children are one level deeper and done).
__global__ void k_bh_moments(BhTree T, int32_t l0, int32_t l1, const uint64_t *__restrict__ keys,
                             const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, const double *__restrict__ m,
                             const double *__restrict__ eps, const double *__restrict__ rbox,
                             double cell0) {
    int i = l0 + blockIdx.x * BHT + threadIdx.x;
    if (i >= l1) return;
    double M = 0.0, sx = 0.0, sy = 0.0, sz = 0.0, em = 0.0;
    if (T.child[i] < 0) {
        for (int q = T.beg[i]; q < T.beg[i] + T.cnt[i]; ++q) {
            double mq = m[q];
            M += mq;
            sx += mq * x[q]; sy += mq * y[q]; sz += mq * z[q];
            em = fmax(em, eps[q]);
        }
    } else {
        for (int c = T.child[i]; c < T.child[i] + T.nchild[i]; ++c) {
            double mc = T.mass[c];
            M += mc;
            sx += mc * T.cx[c]; sy += mc * T.cy[c]; sz += mc * T.cz[c];
            em = fmax(em, T.emax[c]);
        }
    }
    double cx = M > 0.0 ? sx / M : 0.0, cy = M > 0.0 ? sy / M : 0.0, cz = M > 0.0 ? sz / M : 0.0;
    int l = T.level[i];
    double side = cell0 * (double)(1ull << (21 - l));   // cell0 = side of a level-21 cell
    uint64_t pre = l > 0 ? keys[T.beg[i]] >> (3 * (21 - l)) : 0ull;
    double gx = rbox[0] + (compact3(pre >> 2) + 0.5) * side;
    double gy = rbox[2] + (compact3(pre >> 1) + 0.5) * side;
    double gz = rbox[4] + (compact3(pre) + 0.5) * side;
    T.mass[i] = M;
    T.cx[i] = cx; T.cy[i] = cy; T.cz[i] = cz;
    T.gx[i] = gx; T.gy[i] = gy; T.gz[i] = gz;
    T.size[i] = side;
    double dl = sqrt((cx - gx) * (cx - gx) + (cy - gy) * (cy - gy) + (cz - gz) * (cz - gz));
    T.delta[i] = dl;
    T.emax[i] = em;
    BhMono mo;
    mo.cx = cx; mo.cy = cy; mo.cz = cz;
    mo.mass = (float)M;
    mo.emax = (float)em;
    T.mono[i] = mo;
    BhLink lk;
    // rounded up so that the FP32 opening test is never looser than the FP64 one
    lk.size = __double2float_ru(side);
    lk.delta = __double2float_ru(dl);
    lk.child = T.child[i];
    lk.nchild = T.nchild[i];
    lk.beg = T.beg[i];
    lk.cnt = T.cnt[i];
    lk.pad0 = lk.pad1 = 0;
    T.link[i] = lk;
}

// Walk groups: the children (cnt <= 64) of a cell with more than 64
// particles, consecutive ones packed while their sum stays <= 64; the root
// when it has at most 64; runs of 64 inside an unsplittable leaf with more.
// Each particle lands in exactly one group; the list order (atomic append)
// does not affect any particle's result.
__global__ void k_bh_groups(BhTree T, int32_t ncell, int32_t *__restrict__ gbeg,
                            int32_t *__restrict__ gcnt, int32_t *__restrict__ counter) {
    int i = blockIdx.x * BHT + threadIdx.x;
    if (i >= ncell) return;
    int c = T.cnt[i];
    if (i == 0 && c <= 64) {
        int g = atomicAdd(counter, 1);
        gbeg[g] = T.beg[0];
        gcnt[g] = c;
    }
    if (c <= 64) return;
    if (T.child[i] >= 0) {
        // consecutive small children are packed into one group of at most 64
        // (their union stays inside this cell's cube)
        int pb = -1, pc = 0;
        for (int q = T.child[i]; q < T.child[i] + T.nchild[i]; ++q) {
            int cq = T.cnt[q];
            if (cq > 64 || pc + cq > 64) {
                if (pc > 0) {
                    int g = atomicAdd(counter, 1);
                    gbeg[g] = pb;
                    gcnt[g] = pc;
                }
                pb = -1;
                pc = 0;
            }
            if (cq <= 64) {
                if (pc == 0) pb = T.beg[q];
                pc += cq;
            }
        }
        if (pc > 0) {
            int g = atomicAdd(counter, 1);
            gbeg[g] = pb;
            gcnt[g] = pc;
        }
    } else {
        for (int o = 0; o < c; o += 64) {
            int g = atomicAdd(counter, 1);
            gbeg[g] = T.beg[i] + o;
            gcnt[g] = min(64, c - o);
        }
    }
}

// Exclusive prefix sum over the wave of v, 0 <= v < 2^BITS, called with every
// lane active: bit b of the lanes' values is one ballot, and the lanes below
// this one holding it count 2^b each (mbcnt).  Exact integer sums, with no
// LDS round trip.
template <int BITS>
__device__ __forceinline__ int wave_excl_scan(int v, int *total) {
    int pos = 0, tot = 0;
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
        const uint64_t m = __ballot((v >> b) & 1);
        const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        pos += below << b;
        tot += __popcll(m) << b;
    }
    *total = tot;
    return pos;
}

struct BhArgs {
    const double *x, *y, *z, *m, *eps;
    const int32_t *gbeg, *gcnt;   // walk groups
    int64_t n;
    double G, theta;
    double *ax, *ay, *az, *pot;
    int64_t *stats;  // [0] interactions, [1] frontier overflow, [2] max list fill
    int64_t *ovf;    // engine counter of overflowing groups (nullable)
};

// One list entry against lane i, its geometry already formed: (dx, dy, dz)
// from i to the entry, r2 = |d|^2 and t = eps_i + eps_j, twice the pair's
// softening e = (eps_i + eps_j) / 2.  Newtonian where r2 >= t^2 (that is
// r >= 2e: (4e) e and t t round the same product), the cubic-spline softened
// kernel inside.
template <bool POT>
__device__ __forceinline__ void bh_eval(float dx, float dy, float dz, float r2, float t, float mj,
                                        float &ax, float &ay, float &az, float &ph) {
    float f, p;
    // r^-1 once for both branches; inside the softening length it also gives
    // r = r^2 r^-1 (a transcendental costs several FP32 instructions' issue
    // time, and a wave whose lanes split between the branches runs both)
    const float ri = __builtin_amdgcn_rsqf(r2);
    if (r2 >= t * t) {
        f = ri * ri * ri;
        p = ri;
    } else {
        float e = 0.5f * t;
        float r = r2 > 0.0f ? r2 * ri : 0.0f;
        float ie = __builtin_amdgcn_rcpf(e);
        float q = r * ie;
        if (q <= 1.0f) {
            float q2 = q * q;
            f = ie * ie * ie * (4.0f / 3.0f - 1.2f * q2 + 0.5f * q2 * q);
            p = (POT && r2 > 0.0f) ? ie * (1.4f - 2.0f * q2 * (1.0f / 3.0f - 0.15f * q2 + 0.05f * q2 * q))
                                   : 0.0f;
        } else {
            float w = 2.0f - q;
            float w2 = w * w;
            f = ri * ri * ri * (1.0f - w2 * w2 * (1.0f - 0.8f * w + w2 * (1.0f / 6.0f)));
            float q2 = q * q;
            p = POT ? -ri * (1.0f / 15.0f) -
                          ie * q2 * (4.0f / 3.0f - q + 0.3f * q2 - q2 * q * (1.0f / 30.0f)) + 1.6f * ie
                    : 0.0f;
        }
    }
    float s = mj * f;
    ax = __builtin_fmaf(s, dx, ax);
    ay = __builtin_fmaf(s, dy, ay);
    az = __builtin_fmaf(s, dz, az);
    if (POT) ph = __builtin_fmaf(mj, p, ph);
}

// The Newtonian branch of bh_eval alone, for entries every lane sees at
// r >= 2e: the same operations, so the same bits.
template <bool POT>
__device__ __forceinline__ void bh_far(float dx, float dy, float dz, float r2, float mj, float &ax,
                                       float &ay, float &az, float &ph) {
    float ri = __builtin_amdgcn_rsqf(r2);
    float f = ri * ri * ri;
    float s = mj * f;
    ax = __builtin_fmaf(s, dx, ax);
    ay = __builtin_fmaf(s, dy, ay);
    az = __builtin_fmaf(s, dz, az);
    if (POT) ph = __builtin_fmaf(mj, ri, ph);
}

template <bool POT>
__device__ __forceinline__ void bh_pair(float xi, float yi, float zi, float ei, float4 j, float ej,
                                        float &ax, float &ay, float &az, float &ph) {
    float dx = j.x - xi, dy = j.y - yi, dz = j.z - zi;
    float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    bh_eval<POT>(dx, dy, dz, r2, ei + ej, j.w, ax, ay, az, ph);
}

template <bool POT, int SCAP>
__global__ __launch_bounds__(64) void k_bh_walk(BhTree T, BhArgs a) {
    __shared__ int32_t stk[SCAP];
    __shared__ float4 il[BH_ICAP];
    __shared__ float ie[BH_ICAP];
    const int lane = threadIdx.x;
    const int gn = a.gcnt[blockIdx.x];
    const int64_t i = (int64_t)a.gbeg[blockIdx.x] + lane;
    const bool ok = lane < gn;
    double xd = ok ? a.x[i] : 0.0, yd = ok ? a.y[i] : 0.0, zd = ok ? a.z[i] : 0.0;
    double lx = wave_min_d(ok ? xd : INFINITY), hx = wave_max_d(ok ? xd : -INFINITY);
    double ly = wave_min_d(ok ? yd : INFINITY), hy = wave_max_d(ok ? yd : -INFINITY);
    double lz = wave_min_d(ok ? zd : INFINITY), hz = wave_max_d(ok ? zd : -INFINITY);
    const double ox = 0.5 * (lx + hx), oy = 0.5 * (ly + hy), oz = 0.5 * (lz + hz);
    const float xi = (float)(xd - ox), yi = (float)(yd - oy), zi = (float)(zd - oz);
    const float ei = ok ? (float)a.eps[i] : 1.0f;
    float ax = 0.f, ay = 0.f, az = 0.f, ph = 0.f;
    KahanF kx, ky, kz, kp;
    kx.init(); ky.init(); kz.init(); kp.init();
    int icnt = 0;
    int64_t ninter = 0;
    int overflow = 0, maxfill = 0;
    if (lane == 0) stk[0] = 0;
    int sp = 1;
    __syncthreads();

    auto flush = [&]() {
        __syncthreads();
        // entries BH_BATCH at a time, their LDS loads issued before the pairs are
        // evaluated (in list order, as one at a time would)
        int q = 0;
        for (; q + BH_BATCH <= icnt; q += BH_BATCH) {
            float4 j[BH_BATCH];
            float e[BH_BATCH];
#pragma unroll
            for (int k = 0; k < BH_BATCH; ++k) { j[k] = il[q + k]; e[k] = ie[q + k]; }
            float dx[BH_BATCH], dy[BH_BATCH], dz[BH_BATCH], r2[BH_BATCH], t[BH_BATCH];
            bool far = true;
#pragma unroll
            for (int k = 0; k < BH_BATCH; ++k) {
                dx[k] = j[k].x - xi;
                dy[k] = j[k].y - yi;
                dz[k] = j[k].z - zi;
                r2[k] = __builtin_fmaf(dx[k], dx[k], __builtin_fmaf(dy[k], dy[k], dz[k] * dz[k]));
                t[k] = ei + e[k];
                far = far && r2[k] >= t[k] * t[k];
            }
            // lanes past the group's end take part in the vote as far
            if (__all(far || !ok)) {
#pragma unroll
                for (int k = 0; k < BH_BATCH; ++k) bh_far<POT>(dx[k], dy[k], dz[k], r2[k], j[k].w, ax, ay, az, ph);
            } else {
#pragma unroll
                for (int k = 0; k < BH_BATCH; ++k)
                    bh_eval<POT>(dx[k], dy[k], dz[k], r2[k], t[k], j[k].w, ax, ay, az, ph);
            }
        }
        for (; q < icnt; ++q) bh_pair<POT>(xi, yi, zi, ei, il[q], ie[q], ax, ay, az, ph);
        kx.add(ax); ky.add(ay); kz.add(az); kp.add(ph);
        ax = ay = az = ph = 0.f;
        ninter += icnt;
        icnt = 0;
        __syncthreads();
    };

    // Depth-first in batches: pop up to 64 cells from the top of the stack,
    // classify them, push the children of the opened ones.  A popped cell adds
    // at most 7 net entries, so popping at most (cap - slack - sp) / 7 keeps the
    // stack inside cap - slack; past that the walk is pure depth-first (one
    // cell per batch), whose growth is bounded by 7 x the tree depth (<= 147),
    // inside the slack.
    while (sp > 0) {
        int room = (SCAP - BH_SLACK - sp) / 7;
        int take = room < 1 ? 1 : (room < 64 ? room : 64);
        take = sp < take ? sp : take;
        int base = sp - take;
        int node = lane < take ? stk[base + lane] : -1;
        sp = base;
        __syncthreads();
        int act = 0, nent = 0, nfr = 0, lnk = 0;
        BhMono mo{};
        if (node >= 0) {
            mo = T.mono[node];
            const BhLink lk = T.link[node];
            double dx = fmax(0.0, fmax(lx - mo.cx, mo.cx - hx));
            double dy = fmax(0.0, fmax(ly - mo.cy, mo.cy - hy));
            double dz = fmax(0.0, fmax(lz - mo.cz, mo.cz - hz));
            // d theta > s + theta delta, both sides >= 0, compared squared
            const double d2 = dx * dx + dy * dy + dz * dz;
            const double rhs = (double)lk.size + a.theta * (double)lk.delta;
            if (d2 * (a.theta * a.theta) > rhs * rhs) {
                act = 1;
            } else if (lk.child < 0) {
                act = 2;
                nent = lk.cnt;
                lnk = lk.beg;
            } else {
                act = 3;
                nfr = lk.nchild;
                lnk = lk.child;
            }
        }
        // accepted cells: at most one entry per lane
        int tot;
        int pos = wave_excl_scan<1>(act == 1 ? 1 : 0, &tot);
        if (icnt + tot > BH_ICAP) flush();
        if (act == 1) {
            il[icnt + pos] = make_float4((float)(mo.cx - ox), (float)(mo.cy - oy),
                                         (float)(mo.cz - oz), mo.mass);
            ie[icnt + pos] = mo.emax;
        }
        icnt += tot;
        // particles of opened leaves: rounds of at most BH_ROUND per lane, so
        // one round never exceeds 64 * BH_ROUND = BH_ICAP entries.  The lane
        // that opened a leaf only writes its particles' indices into their
        // list slots; the whole wave then loads the particles, one slot per
        // lane (same slots, same list order)
        int rem = act == 2 ? nent : 0;
        int pb = act == 2 ? lnk : 0;
        while (__any(rem > 0)) {
            int tk = rem < BH_ROUND ? rem : BH_ROUND;
            int lpos = wave_excl_scan<3>(tk, &tot);
            if (icnt + tot > BH_ICAP) flush();
            for (int q = 0; q < tk; ++q) ie[icnt + lpos + q] = __int_as_float(pb + q);
            __syncthreads();
            for (int s = lane; s < tot; s += 64) {
                const int p = __float_as_int(ie[icnt + s]);
                il[icnt + s] = make_float4((float)(a.x[p] - ox), (float)(a.y[p] - oy),
                                           (float)(a.z[p] - oz), (float)a.m[p]);
                ie[icnt + s] = (float)a.eps[p];
            }
            __syncthreads();
            icnt += tot;
            rem -= tk;
            pb += tk;
        }
        maxfill = max(maxfill, icnt);
        // children of opened cells onto the stack
        int ftot;
        int fpos = sp + wave_excl_scan<4>(nfr, &ftot);  // 0..8 children
        if (act == 3) {
            int c0 = lnk;
            for (int q = 0; q < nfr; ++q)
                if (fpos + q < SCAP) stk[fpos + q] = c0 + q;
        }
        if (sp + ftot > SCAP) overflow = 1;
        sp = min(sp + ftot, SCAP);
        __syncthreads();
    }
    flush();
    if (ok) {
        a.ax[i] = a.G * kx.value();
        a.ay[i] = a.G * ky.value();
        a.az[i] = a.G * kz.value();
        if (POT) a.pot[i] = -a.G * kp.value();
    }
    if (lane == 0) {
        atomicAdd((unsigned long long *)&a.stats[0], (unsigned long long)(ninter * gn));
        if (overflow) {
            atomicAdd((unsigned long long *)&a.stats[1], 1ull);
            if (a.ovf) atomicAdd((unsigned long long *)a.ovf, 1ull);
        }
        atomicMax((unsigned long long *)&a.stats[2], (unsigned long long)maxfill);
    }
}

}  // namespace

// Host side: tree storage (owned here, not by the engine) and the build.
struct BhHost {
    BhTree T{};
    int64_t cap = 0;
    int32_t *gbeg = nullptr, *gcnt = nullptr;
    int64_t gcap = 0;
    int32_t *nk = nullptr, *off = nullptr, *bounds = nullptr, *total = nullptr, *tsum = nullptr;
    int64_t scratch_cap = 0;
    int64_t *stats = nullptr;
    std::vector<void *> owned;
    ~BhHost() {
        for (void *q : owned) hipFree(q);
    }
    int alloc(void **ptr, size_t bytes) {
        void *q = nullptr;
        if (hipMalloc(&q, bytes ? bytes : 8) != hipSuccess) {
            sphx_set_error("hipMalloc failed for the Barnes-Hut tree");
            return SPHX_ERR_NOMEM;
        }
        owned.push_back(q);
        *ptr = q;
        return SPHX_OK;
    }
};

static int bh_alloc_nodes(BhHost &H, int64_t need) {
    if (need <= H.cap) return SPHX_OK;
    int64_t cap = std::max<int64_t>(need, 2 * H.cap);
    BhTree T{};
    int32_t **ip[] = {&T.beg, &T.cnt, &T.child, &T.nchild, &T.level};
    double **dp[] = {&T.mass, &T.cx, &T.cy, &T.cz, &T.gx, &T.gy, &T.gz, &T.size, &T.delta, &T.emax};
    for (int32_t **q : ip) SPHX_CHECK(H.alloc((void **)q, sizeof(int32_t) * cap));
    for (double **q : dp) SPHX_CHECK(H.alloc((void **)q, sizeof(double) * cap));
    SPHX_CHECK(H.alloc((void **)&T.mono, sizeof(BhMono) * cap));
    SPHX_CHECK(H.alloc((void **)&T.link, sizeof(BhLink) * cap));
    if (H.cap > 0) {
        int32_t *oi[] = {H.T.beg, H.T.cnt, H.T.child, H.T.nchild, H.T.level};
        int32_t *ni[] = {T.beg, T.cnt, T.child, T.nchild, T.level};
        for (int k = 0; k < 5; ++k)
            SPHX_HIP(hipMemcpy(ni[k], oi[k], sizeof(int32_t) * H.cap, hipMemcpyDeviceToDevice));
    }
    H.T = T;
    H.cap = cap;
    return SPHX_OK;
}

// Build the octree and walk it for every particle of the engine (single run,
// Morton-sorted, eps in E.s.eps); accelerations into E.s.gx..gz, potential
// into E.s.pot when with_pot.
int sphx_bh_run(Engine &E, BhHost &H, double theta, bool with_pot, hipStream_t st, BhStats *out) {
    int64_t n = E.n;
    if (n == 0) return SPHX_OK;
    if (E.R != 1 || E.max_runs != 1 || E.rrun >= 0) {
        sphx_set_error("Barnes-Hut gravity needs a single-run engine");
        return SPHX_ERR_STATE;
    }
    const double *eps = E.p.eps_from_h ? E.s.h : E.s.eps;
    hipEvent_t ev[3];
    for (auto &e : ev) SPHX_HIP(hipEventCreate(&e));
    struct EvFree {
        hipEvent_t *e;
        ~EvFree() {
            for (int k = 0; k < 3; ++k) hipEventDestroy(e[k]);
        }
    } evfree{ev};
    SPHX_HIP(hipEventRecord(ev[0], st));
    SPHX_CHECK(bh_alloc_nodes(H, std::max<int64_t>(1024, n / 2)));
    if (!H.total) {
        SPHX_CHECK(H.alloc((void **)&H.total, sizeof(int32_t) * 4));
        SPHX_CHECK(H.alloc((void **)&H.stats, sizeof(int64_t) * 4));
    }
    // root
    int32_t root[5] = {0, (int32_t)n, -1, 0, 0};
    SPHX_HIP(hipMemcpyAsync(H.T.beg, &root[0], 4, hipMemcpyHostToDevice, st));
    SPHX_HIP(hipMemcpyAsync(H.T.cnt, &root[1], 4, hipMemcpyHostToDevice, st));
    SPHX_HIP(hipMemcpyAsync(H.T.level, &root[4], 4, hipMemcpyHostToDevice, st));
    std::vector<int32_t> lev{0, 1};
    int32_t nodes = 1;
    for (int level = 0; level < 21; ++level) {
        int32_t f0 = lev[level], fc = lev[level + 1] - lev[level];
        if (fc <= 0) break;
        if (fc > H.scratch_cap) {
            int64_t c = std::max<int64_t>(fc, 2 * H.scratch_cap);
            SPHX_CHECK(H.alloc((void **)&H.nk, sizeof(int32_t) * c));
            SPHX_CHECK(H.alloc((void **)&H.off, sizeof(int32_t) * c));
            SPHX_CHECK(H.alloc((void **)&H.bounds, sizeof(int32_t) * 9 * c));
            SPHX_CHECK(H.alloc((void **)&H.tsum, sizeof(int32_t) * ((c + BH_SCAN_TILE - 1) / BH_SCAN_TILE)));
            H.scratch_cap = c;
        }
        unsigned g = (unsigned)((fc + BHT - 1) / BHT);
        k_bh_split<<<g, BHT, 0, st>>>(H.T, f0, fc, E.keys, level, H.nk, H.bounds);
        const int ntiles = (int)((fc + BH_SCAN_TILE - 1) / BH_SCAN_TILE);
        k_bh_scan_tiles<<<ntiles, 256, 0, st>>>(H.nk, fc, H.tsum);
        k_bh_scan_sums<<<1, 1024, 0, st>>>(H.tsum, ntiles, H.total);
        k_bh_scan_apply<<<ntiles, 256, 0, st>>>(H.nk, fc, H.tsum, H.off);
        int32_t tot = 0;
        SPHX_HIP(hipMemcpyAsync(&tot, H.total, 4, hipMemcpyDeviceToHost, st));
        SPHX_HIP(hipStreamSynchronize(st));
        if ((int64_t)nodes + tot > H.cap) SPHX_CHECK(bh_alloc_nodes(H, (int64_t)nodes + tot));
        k_bh_emit<<<g, BHT, 0, st>>>(H.T, f0, fc, H.off, H.bounds, nodes, level);
        SPHX_HIP(hipGetLastError());
        nodes += tot;
        lev.push_back(nodes);
        if (tot == 0) break;
    }
    // leaves of the deepest level (never split) get child = -1
    {
        int32_t f0 = lev[lev.size() - 2], fc = lev.back() - f0;
        if (fc > 0 && (int)lev.size() - 2 >= 21) {
            std::vector<int32_t> neg(fc, -1), zero(fc, 0);
            SPHX_HIP(hipMemcpyAsync(H.T.child + f0, neg.data(), 4 * fc, hipMemcpyHostToDevice, st));
            SPHX_HIP(hipMemcpyAsync(H.T.nchild + f0, zero.data(), 4 * fc, hipMemcpyHostToDevice, st));
            SPHX_HIP(hipStreamSynchronize(st));
        }
    }
    int depth = (int)lev.size() - 2;
    // moments bottom-up; the run box gives the cell geometry
    const double *rbox = E.bbox + 6 * (int64_t)E.t.ns;
    double hb[6];
    SPHX_HIP(hipMemcpyAsync(hb, rbox, sizeof(hb), hipMemcpyDeviceToHost, st));
    SPHX_HIP(hipStreamSynchronize(st));
    double ext = std::max(hb[1] - hb[0], std::max(hb[3] - hb[2], hb[5] - hb[4]));
    double cell0 = ext > 0.0 ? ext * (1.0 + 0x1p-30) / (double)(1ull << 21) : 1.0;
    for (int l = depth; l >= 0; --l) {
        int32_t l0 = lev[l], l1 = lev[l + 1];
        if (l1 <= l0) continue;
        k_bh_moments<<<(unsigned)((l1 - l0 + BHT - 1) / BHT), BHT, 0, st>>>(
            H.T, l0, l1, E.keys, E.s.x, E.s.y, E.s.z, E.s.m, eps, rbox, cell0);
    }
    // walk groups
    int64_t gneed = (int64_t)nodes + n / 64 + 1;
    if (gneed > H.gcap) {
        int64_t c = std::max<int64_t>(gneed, 2 * H.gcap);
        SPHX_CHECK(H.alloc((void **)&H.gbeg, sizeof(int32_t) * c));
        SPHX_CHECK(H.alloc((void **)&H.gcnt, sizeof(int32_t) * c));
        H.gcap = c;
    }
    SPHX_HIP(hipMemsetAsync(H.total + 1, 0, sizeof(int32_t), st));
    k_bh_groups<<<(unsigned)((nodes + BHT - 1) / BHT), BHT, 0, st>>>(H.T, nodes, H.gbeg, H.gcnt,
                                                                      H.total + 1);
    SPHX_HIP(hipGetLastError());
    int32_t ngroups = 0;
    SPHX_HIP(hipMemcpyAsync(&ngroups, H.total + 1, 4, hipMemcpyDeviceToHost, st));
    SPHX_HIP(hipStreamSynchronize(st));
    if (ngroups <= 0 || ngroups > gneed) {
        sphx_set_error("Barnes-Hut group list inconsistent with the tree");
        return SPHX_ERR_STATE;
    }
    SPHX_HIP(hipMemsetAsync(H.stats, 0, sizeof(int64_t) * 4, st));
    SPHX_HIP(hipEventRecord(ev[1], st));
    BhArgs a{E.s.x, E.s.y, E.s.z, E.s.m, eps, H.gbeg, H.gcnt, n, E.p.G, theta,
             E.s.gx, E.s.gy, E.s.gz, with_pot ? E.s.pot : nullptr, H.stats,
             E.counters + CNT_BH_OVERFLOW};
    const bool deep = theta < BH_DEEP_THETA;
    if (with_pot && deep) k_bh_walk<true, BH_SCAP_DEEP><<<(unsigned)ngroups, 64, 0, st>>>(H.T, a);
    else if (with_pot) k_bh_walk<true, BH_SCAP><<<(unsigned)ngroups, 64, 0, st>>>(H.T, a);
    else if (deep) k_bh_walk<false, BH_SCAP_DEEP><<<(unsigned)ngroups, 64, 0, st>>>(H.T, a);
    else k_bh_walk<false, BH_SCAP><<<(unsigned)ngroups, 64, 0, st>>>(H.T, a);
    SPHX_HIP(hipGetLastError());
    SPHX_HIP(hipEventRecord(ev[2], st));
    if (out) {
        int64_t sv[4] = {0, 0, 0, 0};
        SPHX_HIP(hipMemcpyAsync(sv, H.stats, sizeof(sv), hipMemcpyDeviceToHost, st));
        SPHX_HIP(hipStreamSynchronize(st));
        out->nodes = nodes;
        out->depth = depth;
        out->interactions = sv[0];
        out->overflow = sv[1];
        out->max_fill = sv[2];
        out->groups = ngroups;
        float b = 0.f, w = 0.f;
        SPHX_HIP(hipEventElapsedTime(&b, ev[0], ev[1]));
        SPHX_HIP(hipEventElapsedTime(&w, ev[1], ev[2]));
        out->build_ms = b;
        out->walk_ms = w;
    }
    return SPHX_OK;
}

BhHost *sphx_bh_host_new() { return new BhHost(); }
void sphx_bh_host_free(BhHost *h) { delete h; }
