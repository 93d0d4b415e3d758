// Direct-sum spline-softened gravity, mixed precision (FP32 pair arithmetic,
// Kahan-compensated FP32 accumulation across j tiles).
//
// Replaces acc_g (11_Dec_2022/GPU_sph/Important_Test_Simulations/
// Evrard_collapse/myCppSPHLibs_v2.h:119-168) and follows the formula of
// getAcc_g_smth (GPU_sph/IC/libsx2_2t.py:633-671); the LDS-tiled layout is the
// pattern of Particles_in_BH_Tree/CUDA_nbody/check.cu:16-93, laid out for
// wave64 and 160 KiB LDS.
//
// Layout: a workgroup of 256 lanes owns an i-block of 256*IPL consecutive
// Morton-ordered particles (IPL per lane, in registers) and streams the run's
// particles through LDS in tiles of TJ as float4 (x, y, z, m) relative to the
// i-block's centre (computed in FP64, rounded once), plus eps as float.
// A j-tile whose AABB is farther than eps_max(I) + eps_max(J) from the
// i-block's AABB has q > 2 for every pair and takes the branch-free 1/r^3
// path (rsq + 12 VALU per pair); the few near tiles take the spline path.
// A far tile whose particles all have the same mass (the usual SPH set-up)
// sums r^-3 d without the mass and scales the tile's partial once (rsq + 11
// VALU per pair).
#include "sphx_internal.h"

namespace {

struct GravTabs {
    const int32_t *ib_beg, *ib_cnt, *ib_run;  // i-blocks
    const int32_t *r_pbeg, *r_pend;           // run particle range
    const int32_t *r_tbeg;                    // run -> first j tile
    const double *tbox;                       // 8 per tile: box(6), emax, unused
    int32_t js;                               // j split factor (>= 1)
};

struct GravIO {
    const double *x, *y, *z, *m, *eps;
    double G;
    double *ax, *ay, *az, *pot;  // js == 1: final outputs
    double *part;                // js > 1: [js][n][4] partial sums
    int64_t n;
    // block time steps: only i-blocks holding a particle whose step ends at
    // substep t1 are computed (blk_B == 0: all)
    const int32_t *bin;
    int32_t blk_B;
    int64_t blk_t1;
};

// Per tile of tj particles: box, max eps, common mass; with bin given (block
// time steps) also whether the tile holds a particle whose step ends at t1.
__global__ void k_tile_boxes(Tables t, const int32_t *__restrict__ r_tbeg, int ntiles_total,
                             int tj, const double *__restrict__ m, double *__restrict__ tbox,
                             const int32_t *__restrict__ bin, int blk_B, int64_t blk_t1,
                             int32_t *__restrict__ act) {
    int tile = blockIdx.x;
    int lane = threadIdx.x;
    // find the run of this tile (runs are few; linear scan by lane 0 is enough)
    __shared__ int srun;
    if (lane == 0) {
        int r = 0;
        while (r + 1 < t.R && r_tbeg[r + 1] <= tile) ++r;
        srun = r;
    }
    __syncthreads();
    int r = srun;
    int k = tile - r_tbeg[r];
    int rb0 = t.s_bbeg[t.r_sbeg[r]];
    int rb1 = t.s_bend[t.r_send[r] - 1];
    int nbt = tj / 64;
    int b0 = rb0 + k * nbt;
    int b1 = min(rb1, b0 + nbt);
    double v[7] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY, 0.0};
    for (int b = b0 + lane; b < b1; b += 64) {
        v[0] = fmin(v[0], t.bxl[b]); v[1] = fmax(v[1], t.bxh[b]);
        v[2] = fmin(v[2], t.byl[b]); v[3] = fmax(v[3], t.byh[b]);
        v[4] = fmin(v[4], t.bzl[b]); v[5] = fmax(v[5], t.bzh[b]);
        v[6] = fmax(v[6], t.bemax[b]);
    }
    for (int a = 0; a < 6; a += 2) {
        v[a] = wave_min_d(v[a]);
        v[a + 1] = wave_max_d(v[a + 1]);
    }
    v[6] = wave_max_d(v[6]);
    // the tile's common mass, or -1 when its masses differ
    int p0 = t.r_pbeg[r] + k * tj, p1 = min(t.r_pend[r], p0 + tj);
    double mlo = INFINITY, mhi = -INFINITY;
    for (int q = p0 + lane; q < p1; q += 64) {
        mlo = fmin(mlo, m[q]);
        mhi = fmax(mhi, m[q]);
    }
    mlo = wave_min_d(mlo);
    mhi = wave_max_d(mhi);
    bool any = false;
    if (bin)
        for (int q = p0 + lane; q < p1; q += 64)
            any = any || (blk_t1 & ((1ll << (blk_B - bin[q])) - 1)) == 0;
    any = __any(any);
    if (lane == 0) {
        for (int a = 0; a < 7; ++a) tbox[8 * tile + a] = v[a];
        tbox[8 * tile + 7] = (mlo == mhi && mlo >= 0.0) ? mlo : -1.0;
        if (bin) act[tile] = any ? 1 : 0;
    }
}

template <int IPL>
struct IState {
    int idx[IPL];  // global index of each i (-1 when the lane holds none)
    float x[IPL], y[IPL], z[IPL], e[IPL];
    float ax[IPL], ay[IPL], az[IPL], ph[IPL];
    KahanF kx[IPL], ky[IPL], kz[IPL], kp[IPL];
};

// UNIF: every j of the tile has the same mass; the mass is left out here and
// applied to the tile's partial sums in flush().
template <int IPL, bool POT, bool UNIF>
__device__ __forceinline__ void far_pair(IState<IPL> &S, float4 j) {
#pragma unroll
    for (int p = 0; p < IPL; ++p) {
        float dx = j.x - S.x[p];
        float dy = j.y - S.y[p];
        float dz = j.z - S.z[p];
        float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        float ri = __builtin_amdgcn_rsqf(r2);
        float mri = UNIF ? ri : j.w * ri;
        float s = mri * (ri * ri);
        S.ax[p] = __builtin_fmaf(s, dx, S.ax[p]);
        S.ay[p] = __builtin_fmaf(s, dy, S.ay[p]);
        S.az[p] = __builtin_fmaf(s, dz, S.az[p]);
        if (POT) S.ph[p] += mri;
    }
}

// Spline-softened pair (q < 2) or point mass (q >= 2).  For 1 < q <= 2 the
// force polynomial of getAcc_g_smth is evaluated in the equivalent form
// 1 - w^4 (1 - 0.8 w + w^2/6), w = 2 - q, which keeps FP32 cancellation small.
template <int IPL, bool POT>
__device__ __forceinline__ void near_pair(IState<IPL> &S, float4 j, float ej, int jidx) {
#pragma unroll
    for (int p = 0; p < IPL; ++p) {
        float dx = j.x - S.x[p];
        float dy = j.y - S.y[p];
        float dz = j.z - S.z[p];
        float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        const float t = S.e[p] + ej;  // 2e; r^2 >= t^2 is r^2 >= (4e) e
        float f, ph;
        // r^-1 once for both branches (as in sym_spline_f)
        const float ri = __builtin_amdgcn_rsqf(r2);
        if (r2 >= t * t) {
            f = ri * ri * ri;
            ph = ri;
        } else {
            const float e = 0.5f * t;
            float r = r2 > 0.0f ? r2 * ri : 0.0f;
            float ie = __builtin_amdgcn_rcpf(e);
            float q = r * ie;
            if (q <= 1.0f) {
                float q2 = q * q;
                f = ie * ie * ie * (4.0f / 3.0f - 1.2f * q2 + 0.5f * q2 * q);
                ph = ie * (1.4f - 2.0f * q2 * (1.0f / 3.0f - 0.15f * q2 + 0.05f * q2 * q));
            } else {
                float w = 2.0f - q;
                float w2 = w * w;
                f = ri * ri * ri * (1.0f - w2 * w2 * (1.0f - 0.8f * w + w2 * (1.0f / 6.0f)));
                float q2 = q * q;
                ph = -ri * (1.0f / 15.0f) -
                     ie * q2 * (4.0f / 3.0f - q + 0.3f * q2 - q2 * q * (1.0f / 30.0f)) + 1.6f * ie;
            }
        }
        float s = j.w * f;
        S.ax[p] = __builtin_fmaf(s, dx, S.ax[p]);
        S.ay[p] = __builtin_fmaf(s, dy, S.ay[p]);
        S.az[p] = __builtin_fmaf(s, dz, S.az[p]);
        // the self pair adds nothing to the force (dx = 0) and is excluded
        // from the potential (j != i in getPE)
        if (POT && jidx != S.idx[p]) S.ph[p] = __builtin_fmaf(j.w, ph, S.ph[p]);
    }
}

// Fold the tile's FP32 partials, scaled by the tile's common mass when the
// tile was summed without it (scale 1 otherwise), into the Kahan sums.
template <int IPL>
__device__ __forceinline__ void flush(IState<IPL> &S, float scale) {
#pragma unroll
    for (int p = 0; p < IPL; ++p) {
        S.kx[p].add(scale * S.ax[p]); S.ax[p] = 0.f;
        S.ky[p].add(scale * S.ay[p]); S.ay[p] = 0.f;
        S.kz[p].add(scale * S.az[p]); S.az[p] = 0.f;
        S.kp[p].add(scale * S.ph[p]); S.ph[p] = 0.f;
    }
}

__device__ __forceinline__ double block_min(double v, double *red) {
    v = wave_min_d(v);
    int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    return fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
}
__device__ __forceinline__ double block_max(double v, double *red) {
    v = wave_max_d(v);
    int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

template <int IPL, int TJ, bool POT>
__global__ __launch_bounds__(256) void k_grav(GravTabs gt, GravIO io) {
    __shared__ float4 jt[TJ];
    __shared__ float je[TJ];
    __shared__ double red[4];
    const int tid = threadIdx.x;
    const int ib = blockIdx.x / gt.js;
    const int jsplit = blockIdx.x % gt.js;
    const int beg = gt.ib_beg[ib], cnt = gt.ib_cnt[ib], run = gt.ib_run[ib];
    const int jbeg = gt.r_pbeg[run], jend = gt.r_pend[run];

    double xd[IPL], yd[IPL], zd[IPL], ed[IPL];
    bool ok[IPL];
    double lx = INFINITY, hx = -INFINITY, ly = INFINITY, hy = -INFINITY, lz = INFINITY,
           hz = -INFINITY, em = 0.0;
#pragma unroll
    for (int p = 0; p < IPL; ++p) {
        int li = tid + p * 256;
        ok[p] = li < cnt;
        int i = beg + (ok[p] ? li : 0);
        xd[p] = io.x[i]; yd[p] = io.y[i]; zd[p] = io.z[i];
        ed[p] = io.eps[i];
        if (ok[p]) {
            lx = fmin(lx, xd[p]); hx = fmax(hx, xd[p]);
            ly = fmin(ly, yd[p]); hy = fmax(hy, yd[p]);
            lz = fmin(lz, zd[p]); hz = fmax(hz, zd[p]);
            em = fmax(em, ed[p]);
        }
    }
    if (io.blk_B > 0) {
        bool any = false;
#pragma unroll
        for (int p = 0; p < IPL; ++p) {
            int b = io.bin[beg + (ok[p] ? tid + p * 256 : 0)];
            any = any || (ok[p] && (io.blk_t1 & ((1ll << (io.blk_B - b)) - 1)) == 0);
        }
        if (!__syncthreads_or(any)) return;  // the same answer in every lane of the block
    }
    lx = block_min(lx, red); hx = block_max(hx, red);
    ly = block_min(ly, red); hy = block_max(hy, red);
    lz = block_min(lz, red); hz = block_max(hz, red);
    em = block_max(em, red);
    const double ox = 0.5 * (lx + hx), oy = 0.5 * (ly + hy), oz = 0.5 * (lz + hz);

    IState<IPL> S;
#pragma unroll
    for (int p = 0; p < IPL; ++p) {
        S.idx[p] = ok[p] ? beg + tid + p * 256 : -1;
        S.x[p] = (float)(xd[p] - ox);
        S.y[p] = (float)(yd[p] - oy);
        S.z[p] = (float)(zd[p] - oz);
        S.e[p] = (float)ed[p];
        S.ax[p] = S.ay[p] = S.az[p] = S.ph[p] = 0.f;
        S.kx[p].init(); S.ky[p].init(); S.kz[p].init(); S.kp[p].init();
    }

    const int ntile = (jend - jbeg + TJ - 1) / TJ;
    const int per = (ntile + gt.js - 1) / gt.js;
    const int t0 = jsplit * per;
    const int t1 = min(ntile, t0 + per);
    for (int tk = t0; tk < t1; ++tk) {
        const int j0 = jbeg + tk * TJ;
        const int jn = min(TJ, jend - j0);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TJ / 256; ++k) {
            int q = tid + k * 256;
            float4 v;
            float e;
            if (q < jn) {
                int j = j0 + q;
                v.x = (float)(io.x[j] - ox);
                v.y = (float)(io.y[j] - oy);
                v.z = (float)(io.z[j] - oz);
                v.w = (float)io.m[j];
                e = (float)io.eps[j];
            } else {
                v.x = 1.0e18f; v.y = 1.0e18f; v.z = 1.0e18f; v.w = 0.0f;
                e = 1.0f;
            }
            jt[q] = v;
            je[q] = e;
        }
        __syncthreads();
        const double *tb = gt.tbox + 8 * (int64_t)(gt.r_tbeg[run] + tk);
        double d2 = sphx_box_d2(lx, hx, ly, hy, lz, hz, tb[0], tb[1], tb[2], tb[3], tb[4], tb[5]);
        double reach = em + tb[6];
        bool far = d2 > reach * reach * (1.0 + 1e-6);
        float scale = 1.0f;
        if (far && tb[7] >= 0.0) {
            scale = (float)tb[7];
#pragma unroll 8
            for (int q = 0; q < TJ; ++q) far_pair<IPL, POT, true>(S, jt[q]);
        } else if (far) {
#pragma unroll 8
            for (int q = 0; q < TJ; ++q) far_pair<IPL, POT, false>(S, jt[q]);
        } else {
            for (int q = 0; q < TJ; ++q) near_pair<IPL, POT>(S, jt[q], je[q], j0 + q);
        }
        flush<IPL>(S, scale);
    }

#pragma unroll
    for (int p = 0; p < IPL; ++p) {
        if (!ok[p]) continue;
        int i = beg + tid + p * 256;
        double ax = io.G * S.kx[p].value();
        double ay = io.G * S.ky[p].value();
        double az = io.G * S.kz[p].value();
        double ph = -io.G * S.kp[p].value();
        if (gt.js == 1) {
            io.ax[i] = ax; io.ay[i] = ay; io.az[i] = az;
            if (POT) io.pot[i] = ph;
        } else {
            double *pp = io.part + ((int64_t)jsplit * io.n + i) * 4;
            pp[0] = ax; pp[1] = ay; pp[2] = az; pp[3] = ph;
        }
    }
}

__global__ void k_grav_reduce(const double *__restrict__ part, int js, int64_t n, int64_t p0,
                              int64_t p1, double *ax, double *ay, double *az, double *pot) {
    int64_t i = p0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p1) return;
    double s[4] = {0, 0, 0, 0};
    for (int k = 0; k < js; ++k) {
        const double *pp = part + ((int64_t)k * n + i) * 4;
        s[0] += pp[0]; s[1] += pp[1]; s[2] += pp[2]; s[3] += pp[3];
    }
    ax[i] = s[0]; ay[i] = s[1]; az[i] = s[2];
    if (pot) pot[i] = s[3];
}

template <int IPL, int TJ>
int launch_grav(const GravTabs &gt, const GravIO &io, int nblocks, bool pot, hipStream_t st) {
    dim3 grid((unsigned)(nblocks * gt.js));
    if (pot) k_grav<IPL, TJ, true><<<grid, 256, 0, st>>>(gt, io);
    else k_grav<IPL, TJ, false><<<grid, 256, 0, st>>>(gt, io);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}

// ---------------------------------------------------------------------------
// Symmetric direct sum (configuration ipl = 8).  Each pair is evaluated once
// and applied to both particles (Newton's third law).  The run's particles
// form blocks of SYM_B; a wave holds one block as its i-set (8 per lane) and
// walks a segment of the upper triangle of block pairs (row I, blocks
// J0..J1-1 >= I).  Per 128-particle chunk of block J, two sets of 64 j
// particles sit one per lane and rotate one lane per step (ds_bpermute, on the
// LDS pipe) together with their own force accumulators, so after 64 steps
// every (i, j) pair of the chunk has been met once and each j is back in its
// lane with its sum over the 512 i.  The j-side sums of row I go to their own
// slots of a partial buffer and the i-side sums of each segment to theirs;
// k_grav_sym_reduce adds them in a fixed order, so results do not depend on
// scheduling.  With tj < SYM_B a segment is a part of one block pair (its
// j particles [c0, c0 + tj)), so that mid-size runs launch several waves per
// block pair.  The diagonal block pair (J = I) is summed one-sided.  Far block
// pairs whose blocks each have one common mass use no mass in the pair and
// scale the chunk sums (rsq + 14 VALU per pair for two interactions).
// ---------------------------------------------------------------------------
constexpr int SYM_B = 512;

struct SymTabs {
    const int32_t *b_beg, *b_cnt;  // blocks (SYM_B particles; a run's last block partial)
    const int32_t *run_b0;         // first block of each run (+ total)
    const int32_t *seg_i, *seg_j0, *seg_j1;  // segments of the block-pair triangle
    const int64_t *row_off;        // per block: first j-side partial of its row
    const int32_t *row_j0;         // per block: first particle after the block in its run
    const int32_t *row_nseg;       // per block: segments of its row
    const int32_t *row_seg0;       // per block: first (global) segment of its row
    const double *bbox;            // per block: box(6), max eps, common mass or -1
    float *px, *py, *pz;           // j-side partials of the current row band
    double *qx, *qy, *qz;          // i-side partials of the band, SYM_B slots per segment
    int32_t seg0;                  // global index of the launch's first segment
    int32_t q_seg0;                // global index of the band's first segment
    int64_t p_base;                // row_off of the band's first row
    const int32_t *act;            // block time steps: blocks holding an active particle
                                   // (nullptr: every block pair is computed)
    const int32_t *run;            // run of each particle
    const int32_t *r_pbeg;         // first particle of each run
    int64_t n;
    const int32_t *order;          // launch order: group g takes segment order[seg0 + g] (global)
    const int32_t *seg_c0;         // per segment: first j of each of its blocks (0, or a
                                   // multiple of jspan when a block pair is split)
    int32_t jspan;                 // j particles a segment takes of each block (SYM_B, 256, 128)
};

// The segment a group takes.  Groups start roughly in index order and the
// last ones set the end of the launch, so the host orders each run's
// segments of a band with the ones nearest the diagonal of the block-pair
// triangle first: those hold the diagonal block and the block pairs within
// softening reach, the longest work, and the short far segments fill the
// tail.  A segment's sums go to its own slots, so the order changes no result.
__device__ __forceinline__ int sym_seg(const SymTabs &T) {
    return T.order[T.seg0 + blockIdx.x] - T.seg0;
}

__device__ __forceinline__ float sym_rot(float v, int addr) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}

// The spline-softened force factor of near_pair, without mass or potential;
// t = eps_i + eps_j is twice the pair's softening e.  Newtonian where
// r^2 >= t^2, the same test as r^2 >= (4e) e (both products round alike).
// r^-1 is taken once, before the softening test, and also gives r = r^2 r^-1
// inside the softening length: a transcendental costs several FP32
// instructions' issue time, and a wave whose lanes split between the two
// branches runs both.
__device__ __forceinline__ float sym_spline_f(float r2, float t) {
    const float ri = __builtin_amdgcn_rsqf(r2);
    if (r2 >= t * t) return ri * ri * ri;
    const float e = 0.5f * t;
    const float r = r2 > 0.0f ? r2 * ri : 0.0f;
    const float ie = __builtin_amdgcn_rcpf(e);
    const float q = r * ie;
    if (q <= 1.0f) {
        float q2 = q * q;
        return ie * ie * ie * (4.0f / 3.0f - 1.2f * q2 + 0.5f * q2 * q);
    }
    float w = 2.0f - q;
    float w2 = w * w;
    return ri * ri * ri * (1.0f - w2 * w2 * (1.0f - 0.8f * w + w2 * (1.0f / 6.0f)));
}

// The i particles a lane holds: P of them (8 on the far path, 4 on the near
// one), with their FP32 force sums of the current j chunk.
template <int P>
struct SymIP {
    float x[P], y[P], z[P], m[P], e[P];
    float ax[P], ay[P], az[P];
};
using SymI = SymIP<8>;

struct SymJ {
    float x, y, z, m, e, fx, fy, fz;
};

// j particle q of a block (lane-relative), relative to the i-block centre;
// positions past the block's end hold massless particles far from any i.
__device__ __forceinline__ SymJ sym_load(const double *x, const double *y, const double *z,
                                         const double *m, const double *eps, int jb, int jc,
                                         int q, double ox, double oy, double oz) {
    SymJ J;
    if (q < jc) {
        int j = jb + q;
        J.x = (float)(x[j] - ox);
        J.y = (float)(y[j] - oy);
        J.z = (float)(z[j] - oz);
        J.m = (float)m[j];
        J.e = (float)eps[j];
    } else {
        J.x = J.y = J.z = -1.0e18f;
        J.m = 0.0f;
        J.e = 1.0f;
    }
    J.fx = J.fy = J.fz = 0.0f;
    return J;
}

// MODE 0: far, common masses (applied later); 1: far; 2: near; 3: diagonal
// block (one-sided, spline).
template <int MODE, int P>
__device__ __forceinline__ void sym_pairs(SymIP<P> &S, SymJ &J) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
        float dx = J.x - S.x[p];
        float dy = J.y - S.y[p];
        float dz = J.z - S.z[p];
        float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        float f;
        if (MODE <= 1) {
            float ri = __builtin_amdgcn_rsqf(r2);
            f = (ri * ri) * ri;
        } else {
            f = sym_spline_f(r2, S.e[p] + J.e);
        }
        float si = MODE == 0 ? f : J.m * f;
        S.ax[p] = __builtin_fmaf(si, dx, S.ax[p]);
        S.ay[p] = __builtin_fmaf(si, dy, S.ay[p]);
        S.az[p] = __builtin_fmaf(si, dz, S.az[p]);
        if (MODE != 3) {
            float sj = MODE == 0 ? f : S.m[p] * f;
            J.fx = __builtin_fmaf(-sj, dx, J.fx);
            J.fy = __builtin_fmaf(-sj, dy, J.fy);
            J.fz = __builtin_fmaf(-sj, dz, J.fz);
        }
    }
}

template <int MODE>
__device__ __forceinline__ void sym_rotate(SymJ &J, int addr) {
    J.x = sym_rot(J.x, addr);
    J.y = sym_rot(J.y, addr);
    J.z = sym_rot(J.z, addr);
    if (MODE != 0) J.m = sym_rot(J.m, addr);
    if (MODE >= 2) J.e = sym_rot(J.e, addr);
    if (MODE != 3) {
        J.fx = sym_rot(J.fx, addr);
        J.fy = sym_rot(J.fy, addr);
        J.fz = sym_rot(J.fz, addr);
    }
}

// Near-path pairs with a far mask: bit p set means every pair of i slot p
// with this j set is beyond the softening reach, so the spline's r^-3 branch
// (the same operations) is taken without its test.  far is wave-uniform.
template <int P>
__device__ __forceinline__ void sym_pairs_masked(SymIP<P> &S, SymJ &J, unsigned far) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
        float dx = J.x - S.x[p];
        float dy = J.y - S.y[p];
        float dz = J.z - S.z[p];
        float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        float f;
        if (far & (1u << p)) {
            float ri = __builtin_amdgcn_rsqf(r2);
            f = (ri * ri) * ri;
        } else {
            f = sym_spline_f(r2, S.e[p] + J.e);
        }
        float si = J.m * f;
        S.ax[p] = __builtin_fmaf(si, dx, S.ax[p]);
        S.ay[p] = __builtin_fmaf(si, dy, S.ay[p]);
        S.az[p] = __builtin_fmaf(si, dz, S.az[p]);
        float sj = S.m[p] * f;
        J.fx = __builtin_fmaf(-sj, dx, J.fx);
        J.fy = __builtin_fmaf(-sj, dy, J.fy);
        J.fz = __builtin_fmaf(-sj, dz, J.fz);
    }
}

template <int P>
__device__ __forceinline__ void sym_chunk_masked(SymIP<P> &S, SymJ &A, SymJ &B, int addr,
                                                 unsigned farA, unsigned farB) {
    for (int s = 0; s < 64; ++s) {
        sym_pairs_masked(S, A, farA);
        sym_rotate<2>(A, addr);
        sym_pairs_masked(S, B, farB);
        sym_rotate<2>(B, addr);
    }
}

template <int MODE, int P = 8>
__device__ __forceinline__ void sym_chunk(SymIP<P> &S, SymJ &A, SymJ &B, int addr) {
    for (int s = 0; s < 64; ++s) {
        sym_pairs<MODE>(S, A);
        sym_rotate<MODE>(A, addr);
        sym_pairs<MODE>(S, B);
        sym_rotate<MODE>(B, addr);
    }
}

// Block-pair mode of tile J for row I (wave-uniform): 3 diagonal, 2 near
// (spline path), 0 far with common masses on both sides, 1 far.
__device__ __forceinline__ int sym_mode(const double *bI, const double *bJ, bool diag) {
    if (diag) return 3;
    double d2 = sphx_box_d2(bI[0], bI[1], bI[2], bI[3], bI[4], bI[5], bJ[0], bJ[1], bJ[2], bJ[3],
                            bJ[4], bJ[5]);
    double reach = bI[6] + bJ[6];
    bool far = d2 > reach * reach * (1.0 + 1e-6);
    return far ? ((bI[7] >= 0.0 && bJ[7] >= 0.0) ? 0 : 1) : 2;
}

// One segment of row I, restricted to its far block pairs: the i block held
// as 8 per lane.  Its i-side sums are stored in the segment's slots, or added
// to what the segment's near pairs stored there before (add).  Every j-side
// slot belongs to exactly one block pair, near or far.
__device__ __forceinline__ void sym_segment_far(const SymTabs &T, const double *__restrict__ x,
                                                const double *__restrict__ y,
                                                const double *__restrict__ z,
                                                const double *__restrict__ m,
                                                const double *__restrict__ eps, bool add) {
    const int lane = threadIdx.x;
    const int sg = sym_seg(T);
    const int I = T.seg_i[sg], J0 = T.seg_j0[sg], J1 = T.seg_j1[sg];
    const int beg = T.b_beg[I], cnt = T.b_cnt[I];
    const double *bI = T.bbox + 8 * (int64_t)I;
    // block time steps: a block pair with no active particle on either side is
    // not needed (its sums feed only inactive particles)
    const bool actI = !T.act || T.act[I];
    const double ox = 0.5 * (bI[0] + bI[1]), oy = 0.5 * (bI[2] + bI[3]), oz = 0.5 * (bI[4] + bI[5]);
    const double mI = bI[7];
    const int addr = ((lane + 1) & 63) << 2;  // lane L takes the value of lane L + 1

    SymI S;
    KahanF kx[8], ky[8], kz[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
        int li = lane + 64 * p;
        bool ok = li < cnt;
        int i = beg + (ok ? li : 0);
        S.x[p] = ok ? (float)(x[i] - ox) : 1.0e18f;
        S.y[p] = ok ? (float)(y[i] - oy) : 1.0e18f;
        S.z[p] = ok ? (float)(z[i] - oz) : 1.0e18f;
        S.m[p] = ok ? (float)m[i] : 0.0f;
        S.e[p] = ok ? (float)eps[i] : 1.0f;
        S.ax[p] = S.ay[p] = S.az[p] = 0.0f;
        kx[p].init(); ky[p].init(); kz[p].init();
    }
    const int64_t roff = T.row_off[I] - T.p_base;
    const int rj0 = T.row_j0[I];
    const int c0 = T.seg_c0[sg];

    for (int J = J0; J < J1; ++J) {
        const int jb = T.b_beg[J], jc = T.b_cnt[J];
        const double *bJ = T.bbox + 8 * (int64_t)J;
        if (!actI && !T.act[J]) continue;
        const int mode = sym_mode(bI, bJ, J == I);
        if (mode > 1) continue;
        const float si = mode == 0 ? (float)bJ[7] : 1.0f;  // i-side scale: the j block's mass
        const float sj = mode == 0 ? (float)mI : 1.0f;     // j-side scale: the i block's mass
        const int ce = jc < c0 + T.jspan ? jc : c0 + T.jspan;
        for (int c = c0; c < ce; c += 128) {
            SymJ A = sym_load(x, y, z, m, eps, jb, jc, c + lane, ox, oy, oz);
            SymJ B = sym_load(x, y, z, m, eps, jb, jc, c + 64 + lane, ox, oy, oz);
            if (mode == 0) sym_chunk<0>(S, A, B, addr);
            else sym_chunk<1>(S, A, B, addr);
            int qa = c + lane, qb = c + 64 + lane;
            if (qa < jc) {
                int64_t o = roff + (jb + qa - rj0);
                T.px[o] = sj * A.fx; T.py[o] = sj * A.fy; T.pz[o] = sj * A.fz;
            }
            if (qb < jc) {
                int64_t o = roff + (jb + qb - rj0);
                T.px[o] = sj * B.fx; T.py[o] = sj * B.fy; T.pz[o] = sj * B.fz;
            }
#pragma unroll
            for (int p = 0; p < 8; ++p) {
                kx[p].add(si * S.ax[p]); S.ax[p] = 0.0f;
                ky[p].add(si * S.ay[p]); S.ay[p] = 0.0f;
                kz[p].add(si * S.az[p]); S.az[p] = 0.0f;
            }
        }
    }
#pragma unroll
    for (int p = 0; p < 8; ++p) {
        int li = lane + 64 * p;
        if (li < cnt) {
            int64_t o = (int64_t)(T.seg0 + sg - T.q_seg0) * SYM_B + li;
            if (add) {
                T.qx[o] += kx[p].value();
                T.qy[o] += ky[p].value();
                T.qz[o] += kz[p].value();
            } else {
                T.qx[o] = kx[p].value();
                T.qy[o] = ky[p].value();
                T.qz[o] = kz[p].value();
            }
        }
    }
}

// The same segment restricted to its near and diagonal block pairs.  The
// block's 512 i are taken as two halves of 256 (4 per lane), one after the
// other: the registers of this path then stay within the far path's budget.
// Each of a half's four i slots (64 particles, Morton-contiguous) is tested
// against each 64-particle j set of a chunk, so that the pairs beyond the
// softening reach take the r^-3 path.  The j sets rotate once per half; the
// first half stores their j-side sums and the second adds to them (lane L
// owns the same j slots in both).  Stores the segment's i-side sums and
// returns true, or returns false at once when the segment has no near pair.
__device__ __forceinline__ bool sym_segment_near(const SymTabs &T, const double *__restrict__ x,
                                                 const double *__restrict__ y,
                                                 const double *__restrict__ z,
                                                 const double *__restrict__ m,
                                                 const double *__restrict__ eps) {
    const int lane = threadIdx.x;
    const int sg = sym_seg(T);
    const int I = T.seg_i[sg], J0 = T.seg_j0[sg], J1 = T.seg_j1[sg];
    const int beg = T.b_beg[I], cnt = T.b_cnt[I];
    const double *bI = T.bbox + 8 * (int64_t)I;
    const bool actI = !T.act || T.act[I];
    // most segments hold no near pair: leave before loading the i block
    bool any = false;
    for (int J = J0; J < J1 && !any; ++J)
        any = (actI || T.act[J]) && sym_mode(bI, T.bbox + 8 * (int64_t)J, J == I) >= 2;
    if (!any) return false;
    const double ox = 0.5 * (bI[0] + bI[1]), oy = 0.5 * (bI[2] + bI[3]), oz = 0.5 * (bI[4] + bI[5]);
    const int addr = ((lane + 1) & 63) << 2;  // lane L takes the value of lane L + 1
    const int64_t roff = T.row_off[I] - T.p_base;
    const int rj0 = T.row_j0[I];
    const int c0 = T.seg_c0[sg];
    __shared__ float slotbox[4][7];

    for (int h = 0; 256 * h < cnt; ++h) {
        SymIP<4> S;
        KahanF kx[4], ky[4], kz[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            int li = lane + 64 * (4 * h + p);
            bool ok = li < cnt;
            int i = beg + (ok ? li : 0);
            S.x[p] = ok ? (float)(x[i] - ox) : 1.0e18f;
            S.y[p] = ok ? (float)(y[i] - oy) : 1.0e18f;
            S.z[p] = ok ? (float)(z[i] - oz) : 1.0e18f;
            S.m[p] = ok ? (float)m[i] : 0.0f;
            S.e[p] = ok ? (float)eps[i] : 1.0f;
            S.ax[p] = S.ay[p] = S.az[p] = 0.0f;
            kx[p].init(); ky[p].init(); kz[p].init();
        }
        // the box (block-relative FP32 coordinates: min and max are exact) and
        // largest softening of each slot p, the 64 i particles lane + 64 (4h + p),
        // kept in LDS for the chunk tests
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            bool ok = lane + 64 * (4 * h + p) < cnt;
            float v[7] = {wave_min_f(ok ? S.x[p] : INFINITY), wave_max_f(ok ? S.x[p] : -INFINITY),
                          wave_min_f(ok ? S.y[p] : INFINITY), wave_max_f(ok ? S.y[p] : -INFINITY),
                          wave_min_f(ok ? S.z[p] : INFINITY), wave_max_f(ok ? S.z[p] : -INFINITY),
                          wave_max_f(ok ? S.e[p] : 0.0f)};
            if (lane < 7) slotbox[p][lane] = v[lane];
        }
        __syncthreads();

        for (int J = J0; J < J1; ++J) {
            const int jb = T.b_beg[J], jc = T.b_cnt[J];
            const double *bJ = T.bbox + 8 * (int64_t)J;
            if (!actI && !T.act[J]) continue;
            const int mode = sym_mode(bI, bJ, J == I);
            if (mode <= 1) continue;
            const int ce = jc < c0 + T.jspan ? jc : c0 + T.jspan;
            for (int c = c0; c < ce; c += 128) {
                SymJ A = sym_load(x, y, z, m, eps, jb, jc, c + lane, ox, oy, oz);
                SymJ B = sym_load(x, y, z, m, eps, jb, jc, c + 64 + lane, ox, oy, oz);
                int cm = mode;
                unsigned farA = 0, farB = 0;
                if (mode == 2) {
                    // A near block pair still holds many far pairs: test each i slot's
                    // box against each j set's (both from the FP32 relative
                    // coordinates the pairs use, with a margin for their rounding).
                    // A (slot, set) beyond reach takes r^-3 without the spline's test;
                    // the forces are the same, as that test would pass for all its pairs.
                    // lane q < 4 tests slot q against set A, lane 4 + q against set B
                    const bool va = c + lane < jc, vb = c + 64 + lane < jc;
                    float sv[14] = {wave_min_f(va ? A.x : INFINITY), wave_max_f(va ? A.x : -INFINITY),
                                    wave_min_f(va ? A.y : INFINITY), wave_max_f(va ? A.y : -INFINITY),
                                    wave_min_f(va ? A.z : INFINITY), wave_max_f(va ? A.z : -INFINITY),
                                    wave_max_f(va ? A.e : 0.0f),
                                    wave_min_f(vb ? B.x : INFINITY), wave_max_f(vb ? B.x : -INFINITY),
                                    wave_min_f(vb ? B.y : INFINITY), wave_max_f(vb ? B.y : -INFINITY),
                                    wave_min_f(vb ? B.z : INFINITY), wave_max_f(vb ? B.z : -INFINITY),
                                    wave_max_f(vb ? B.e : 0.0f)};
                    bool far_q = false;
                    if (lane < 8) {
                        const float *jb7 = lane < 4 ? sv : sv + 7;
                        const float *ib7 = slotbox[lane & 3];
                        double big = 1.0;
                        for (int k = 0; k < 6; ++k)
                            big = fmax(big, fmax(fabs((double)jb7[k]), fabs((double)ib7[k])));
                        double d2 = sphx_box_d2(ib7[0], ib7[1], ib7[2], ib7[3], ib7[4], ib7[5],
                                                jb7[0], jb7[1], jb7[2], jb7[3], jb7[4], jb7[5]);
                        double reach = ((double)ib7[6] + (double)jb7[6]) * (1.0 + 1e-5) + 4e-6 * big;
                        far_q = d2 > reach * reach;
                    }
                    uint64_t fm = __ballot(far_q);
                    farA = (unsigned)(fm & 15u);
                    farB = (unsigned)((fm >> 4) & 15u);
                    if (farA == 15u && farB == 15u) cm = 1;
                }
                if (cm == 1) sym_chunk<1, 4>(S, A, B, addr);
                else if (cm == 2 && (farA | farB) == 0) sym_chunk<2, 4>(S, A, B, addr);
                else if (cm == 2) sym_chunk_masked<4>(S, A, B, addr, farA, farB);
                else sym_chunk<3, 4>(S, A, B, addr);
                if (mode != 3) {
                    int qa = c + lane, qb = c + 64 + lane;
                    if (qa < jc) {
                        int64_t o = roff + (jb + qa - rj0);
                        if (h == 0) {
                            T.px[o] = A.fx; T.py[o] = A.fy; T.pz[o] = A.fz;
                        } else {
                            T.px[o] += A.fx; T.py[o] += A.fy; T.pz[o] += A.fz;
                        }
                    }
                    if (qb < jc) {
                        int64_t o = roff + (jb + qb - rj0);
                        if (h == 0) {
                            T.px[o] = B.fx; T.py[o] = B.fy; T.pz[o] = B.fz;
                        } else {
                            T.px[o] += B.fx; T.py[o] += B.fy; T.pz[o] += B.fz;
                        }
                    }
                }
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    kx[p].add(S.ax[p]); S.ax[p] = 0.0f;
                    ky[p].add(S.ay[p]); S.ay[p] = 0.0f;
                    kz[p].add(S.az[p]); S.az[p] = 0.0f;
                }
            }
        }
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            int li = lane + 64 * (4 * h + p);
            if (li < cnt) {
                int64_t o = (int64_t)(T.seg0 + sg - T.q_seg0) * SYM_B + li;
                T.qx[o] = kx[p].value();
                T.qy[o] = ky[p].value();
                T.qz[o] = kz[p].value();
            }
        }
    }
    return true;
}

// One wave per segment: its near block pairs, then its far ones (three waves
// per SIMD: both paths fit 168 VGPRs without spills).  Near work, a few
// percent of the pairs at 1M particles but a third at 65K, is spread over the
// same launch as the far work instead of trailing it.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_grav_sym(
    SymTabs T, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ m, const double *__restrict__ eps) {
    const bool near = sym_segment_near(T, x, y, z, m, eps);
    sym_segment_far(T, x, y, z, m, eps, near);
}


// The block-pair triangle is summed in bands of rows [lo, hi), sized so that a
// band's partial sums fit a memory budget.  For particle i of block I (row I):
//   a_i = G (sum over the earlier rows of the run of their j-side sums for i,
//            in row order, + sum over the segments of row I of their i-side
//            sums, in order),
// one FP64 accumulator per component, held in g between bands (zeroed before
// the first band by k_grav_sym_zero).  Each band adds the terms of its rows in
// that order and finishes the particles whose own row it holds, so the sums do
// not depend on how the rows are banded.
__global__ void k_grav_sym_zero(int64_t p0, int64_t p1, double *gx, double *gy, double *gz) {
    int64_t i = p0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < p1) gx[i] = gy[i] = gz[i] = 0.0;
}

__global__ void k_grav_sym_reduce(SymTabs T, double G, int64_t p0, int64_t p1, int lo, int hi,
                                  double *gx, double *gy, double *gz) {
    int64_t i = p0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p1) return;
    int r = T.run[i];
    int b0 = T.run_b0[r];
    int I = b0 + (int)((i - T.r_pbeg[r]) / SYM_B);
    if (I < lo) return;                            // finished by an earlier band
    const int a0 = b0 > lo ? b0 : lo, a1 = I < hi ? I : hi;
    const bool own = I < hi;
    if (a1 <= a0 && !own) return;                  // no row of this band reaches i
    double sx = gx[i], sy = gy[i], sz = gz[i];
    // unrolled so that the loads of several slots are in flight at once (the
    // additions keep their order)
#pragma unroll 8
    for (int I2 = a0; I2 < a1; ++I2) {
        int64_t o = T.row_off[I2] - T.p_base + (i - T.row_j0[I2]);
        sx += (double)T.px[o]; sy += (double)T.py[o]; sz += (double)T.pz[o];
    }
    if (own) {
        int ns = T.row_nseg[I], sq = T.row_seg0[I] - T.q_seg0;
        int64_t li = i - T.b_beg[I];
#pragma unroll 8
        for (int k = 0; k < ns; ++k) {
            int64_t o = (int64_t)(sq + k) * SYM_B + li;
            sx += T.qx[o]; sy += T.qy[o]; sz += T.qz[o];
        }
        sx = G * sx; sy = G * sy; sz = G * sz;
    }
    gx[i] = sx;
    gy[i] = sy;
    gz[i] = sz;
}

// Issue-rate reference for the far block pairs of k_grav_sym: the same
// instruction mix per pair (3 sub, r^2 as 1 mul + 2 fma, rsq, 2 mul, 3 fma to
// the i-side sums, 3 fma to the j-side sums) and the same shape (8 i and 2
// sets of j per lane), on register-resident particles with no memory or LDS
// traffic.  Each step moves every j by s so that nothing is loop-invariant.
// Each lane evaluates 16 pairs per step; one store per lane at the end.
__global__ __launch_bounds__(256) void k_pair_probe(float *out, int iters, float s) {
    float xi[8], yi[8], zi[8], ax[8], ay[8], az[8];
    const float t = threadIdx.x * 1e-3f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        xi[k] = t + k;
        yi[k] = t - k;
        zi[k] = 0.5f * k;
        ax[k] = ay[k] = az[k] = 0.f;
    }
    float xj[2], yj[2], zj[2], bx[2], by[2], bz[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        xj[q] = 10.f + q;
        yj[q] = 11.f;
        zj[q] = 12.f - q;
        bx[q] = by[q] = bz[q] = 0.f;
    }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const float dx = xj[q] - xi[k], dy = yj[q] - yi[k], dz = zj[q] - zi[k];
                const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float inv3 = inv * inv * inv;
                ax[k] = __builtin_fmaf(inv3, dx, ax[k]);
                ay[k] = __builtin_fmaf(inv3, dy, ay[k]);
                az[k] = __builtin_fmaf(inv3, dz, az[k]);
                bx[q] = __builtin_fmaf(-inv3, dx, bx[q]);
                by[q] = __builtin_fmaf(-inv3, dy, by[q]);
                bz[q] = __builtin_fmaf(-inv3, dz, bz[q]);
            }
            xj[q] += s;
            yj[q] += s;
            zj[q] += s;
        }
    }
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += ax[k] + ay[k] + az[k];
#pragma unroll
    for (int q = 0; q < 2; ++q) acc += bx[q] + by[q] + bz[q];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

}  // namespace

// Host tables of the symmetric sum; rebuilt when the run sizes or the segment
// length change.
struct SymHost {
    int seg_len = 0;
    int seg_tj = 0;                 // j particles per segment and block (SYM_B: whole blocks)
    std::vector<int64_t> sizes;
    int nblocks = 0, nseg = 0;
    std::vector<int32_t> run_seg0;  // first segment of each run (+ total)
    int32_t *dev_i = nullptr;       // int32 tables
    int64_t dev_i_cap = 0;
    int64_t *row_off = nullptr;
    int64_t row_off_cap = 0;
    double *bbox = nullptr;
    int64_t bbox_cap = 0;
    float *part = nullptr;          // 3 x total row length
    int64_t part_cap = 0, part_len = 0;
    double *q = nullptr;            // 3 x SYM_B per segment
    int64_t q_cap = 0;
    int32_t *act = nullptr;         // per block: holds an active particle (block time steps)
    int64_t act_cap = 0;
    int64_t budget = -1;            // bytes reserved for a band's partial sums (-1: none)
    double *pool = nullptr;         // reserved memory of a budget (i-side, then j-side sums)
    // row bands of the block-pair triangle whose partial sums fit the budget
    // (or SYM_BAND_BYTES without one)
    struct Band {
        int row_lo, row_hi, seg_lo, seg_hi;
        int64_t p_base, slots;
    };
    std::vector<Band> bands;
    int o_bbeg = 0, o_bcnt = 0, o_runb0 = 0, o_si = 0, o_sj0 = 0, o_sj1 = 0, o_rj0 = 0, o_rns = 0,
        o_rs0 = 0, o_ord = 0, o_c0 = 0;
};

void sphx_sym_host_free(SymHost *h) { delete h; }

static int64_t sym_blocks(int64_t n) { return (n + SYM_B - 1) / SYM_B; }

// Partial-sum memory of one band when no budget is reserved (plans and
// stateless calls): 16 GiB holds N = 1,048,576 in one band; larger runs are
// summed in several.
constexpr int64_t SYM_BAND_BYTES = int64_t(16) << 30;

// Bytes of partial sums of one row of the triangle: 12 per j-side slot (row a
// of nb blocks holds the particles after block a) and 24 x SYM_B per segment.
static int64_t sym_row_bytes(int64_t slots, int64_t segs) { return 12 * slots + 24 * SYM_B * segs; }

// Bytes of partial sums one run of n particles needs at segment length S, all
// rows together.
int64_t sphx_sym_run_bytes(int64_t n, int S) {
    int64_t nb = sym_blocks(n), bytes = 0;
    for (int64_t a = 0; a < nb; ++a)
        bytes += sym_row_bytes(n - std::min(n, (a + 1) * SYM_B), (nb - a + S - 1) / S);
    return bytes;
}

// The least band memory with which a run of n particles can be summed: its
// first (largest) row.
int64_t sphx_sym_min_bytes(int64_t n, int S) {
    if (n <= 0) return 0;
    int64_t nb = sym_blocks(n);
    return sym_row_bytes(n - std::min<int64_t>(n, SYM_B), (nb + S - 1) / S);
}

// Fix the partial-sum memory of an engine (served batches): allocated once;
// the rows of every run are summed in bands that fit it, and runs whose
// largest row would not fit are refused at admission.
int sphx_sym_reserve(Engine &E, int S, int64_t budget_bytes) {
    if (!E.grav.sym) E.grav.sym = new SymHost();
    SymHost &H = *E.grav.sym;
    if (H.budget >= 0) { sphx_set_error("symmetric gravity memory already reserved"); return SPHX_ERR_STATE; }
    SPHX_CHECK(E.alloc((void **)&H.pool, (size_t)std::max<int64_t>(budget_bytes, 64)));
    H.budget = budget_bytes;
    H.seg_len = 0;  // rebuild the tables at the next launch
    (void)S;
    return SPHX_OK;
}

// Whether the runs `sizes` fit the reserved partial-sum memory at segment
// length S (always true without a reservation).
bool sphx_sym_fits(const Engine &E, const std::vector<int64_t> &sizes, int S) {
    if (!E.grav.sym || E.grav.sym->budget < 0) return true;
    for (int64_t n : sizes)
        if (sphx_sym_min_bytes(n, S) > E.grav.sym->budget) return false;
    return true;
}

static int sym_prepare(Engine &E, const GravConfig &cfg, hipStream_t st) {
    if (!E.grav.sym) E.grav.sym = new SymHost();
    SymHost &H = *E.grav.sym;
    const int S = cfg.js;
    // tj < SYM_B (with one block pair per segment) splits every block pair
    // into SYM_B / tj segments of tj j particles each: more, shorter waves
    const int span = cfg.tj < SYM_B ? cfg.tj : SYM_B;
    if (H.seg_len == S && H.seg_tj == span && H.sizes == E.run_n) return SPHX_OK;
    std::vector<int32_t> bbeg, bcnt, runb0, si, sj0, sj1, sc0, rj0, rns, rs0;
    std::vector<int64_t> roff;
    std::vector<int32_t> run_seg0;
    int64_t off = 0, plen = 0;
    for (int r = 0; r < E.R; ++r) {
        int64_t nr = E.run_n[r];
        int b0 = (int)bbeg.size();
        runb0.push_back(b0);
        run_seg0.push_back((int32_t)si.size());
        for (int64_t a = 0; a < nr; a += SYM_B) {
            bbeg.push_back((int32_t)(off + a));
            bcnt.push_back((int32_t)std::min<int64_t>(SYM_B, nr - a));
        }
        int b1 = (int)bbeg.size();
        for (int I = b0; I < b1; ++I) {
            int64_t j0 = (int64_t)bbeg[I] + bcnt[I];  // first particle after block I
            rj0.push_back((int32_t)j0);
            roff.push_back(plen);
            plen += (off + nr) - j0;
            rs0.push_back((int32_t)si.size());
            int k = 0;
            for (int J = I; J < b1; J += S) {
                for (int c = 0; c < bcnt[J]; c += span, ++k) {
                    si.push_back(I);
                    sj0.push_back(J);
                    sj1.push_back(std::min(b1, J + S));
                    sc0.push_back(c);
                }
            }
            rns.push_back(k);
        }
        off += nr;
    }
    runb0.push_back((int32_t)bbeg.size());
    run_seg0.push_back((int32_t)si.size());
    int nb = (int)bbeg.size(), ns = (int)si.size();
    // row bands: consecutive rows while their partial sums fit the budget
    const int64_t band_bytes = H.budget >= 0 ? H.budget : SYM_BAND_BYTES;
    std::vector<SymHost::Band> bands;
    int64_t max_slots = 0, max_segs = 0;
    for (int I = 0; I < nb;) {
        SymHost::Band bd{I, I, rs0[I], rs0[I], roff[I], 0};
        int64_t segs = 0;
        while (I < nb) {
            int64_t rsl = (I + 1 < nb ? roff[I + 1] : plen) - roff[I];
            if (sym_row_bytes(bd.slots + rsl, segs + rns[I]) > band_bytes) {
                if (I == bd.row_lo) {
                    sphx_set_error("a row of the symmetric gravity sum (" +
                                   std::to_string(sym_row_bytes(rsl, rns[I])) +
                                   " bytes of partial sums) exceeds its reserved memory");
                    return SPHX_ERR_STATE;
                }
                break;
            }
            bd.slots += rsl;
            segs += rns[I];
            ++I;
        }
        bd.row_hi = I;
        bd.seg_hi = I < nb ? rs0[I] : ns;
        max_slots = std::max(max_slots, bd.slots);
        max_segs = std::max(max_segs, segs);
        bands.push_back(bd);
    }
    int64_t need_p = 3 * max_slots, need_q = 3 * (int64_t)SYM_B * max_segs;
    // launch order (sym_seg): within each run's part of each band, segments
    // by their distance from the diagonal, nearest first: the k-th segment of
    // every row of the part (first block pair J0 = I + k S; with split block
    // pairs, part k mod (SYM_B / tj) of block I + k / (SYM_B / tj)), for k = 0, 1, ...,
    // rows in order.  The rows of a part are consecutive and their segments
    // contiguous, so this is a permutation of the part's range, built in one
    // pass (the tables are rebuilt whenever the run set changes).
    std::vector<int32_t> ord;
    ord.reserve(ns);
    for (const SymHost::Band &bd : bands) {
        for (int r = 0; r < E.R; ++r) {
            const int r0 = std::max(bd.row_lo, runb0[r]), r1 = std::min(bd.row_hi, runb0[r + 1]);
            int kmax = 0;
            for (int I = r0; I < r1; ++I) kmax = std::max(kmax, rns[I]);
            for (int k = 0; k < kmax; ++k)
                for (int I = r0; I < r1; ++I)
                    if (k < rns[I]) ord.push_back(rs0[I] + k);
        }
    }
    if ((int)ord.size() != ns) {
        sphx_set_error("symmetric gravity: segment order does not cover the segments");
        return SPHX_ERR_STATE;
    }
    std::vector<int32_t> all;
    auto put = [&](const std::vector<int32_t> &v) {
        int o = (int)all.size();
        all.insert(all.end(), v.begin(), v.end());
        return o;
    };
    int o_bbeg = put(bbeg), o_bcnt = put(bcnt), o_runb0 = put(runb0), o_si = put(si),
        o_sj0 = put(sj0), o_sj1 = put(sj1), o_rj0 = put(rj0), o_rns = put(rns), o_rs0 = put(rs0),
        o_ord = put(ord), o_c0 = put(sc0);
    if ((int64_t)all.size() > H.dev_i_cap) {
        SPHX_CHECK(E.alloc((void **)&H.dev_i, sizeof(int32_t) * (all.size() + 1024)));
        H.dev_i_cap = (int64_t)all.size() + 1024;
    }
    if (nb > H.row_off_cap) {
        SPHX_CHECK(E.alloc((void **)&H.row_off, sizeof(int64_t) * (nb + 64)));
        H.row_off_cap = nb + 64;
    }
    if (8 * (int64_t)nb > H.bbox_cap) {
        SPHX_CHECK(E.alloc((void **)&H.bbox, sizeof(double) * (8 * (int64_t)nb + 64)));
        H.bbox_cap = 8 * (int64_t)nb + 64;
    }
    if (H.budget >= 0) {
        H.q = H.pool;  // carved from the reservation band by band (sym_run)
        H.part = nullptr;
    } else {
        if (need_p > H.part_cap) {
            SPHX_CHECK(E.alloc((void **)&H.part, sizeof(float) * (need_p + 64)));
            H.part_cap = need_p + 64;
        }
        if (need_q > H.q_cap) {
            SPHX_CHECK(E.alloc((void **)&H.q, sizeof(double) * (need_q + 64)));
            H.q_cap = need_q + 64;
        }
    }
    if (!all.empty())
        SPHX_HIP(hipMemcpyAsync(H.dev_i, all.data(), sizeof(int32_t) * all.size(), hipMemcpyHostToDevice, st));
    if (nb > 0)
        SPHX_HIP(hipMemcpyAsync(H.row_off, roff.data(), sizeof(int64_t) * nb, hipMemcpyHostToDevice, st));
    SPHX_HIP(hipStreamSynchronize(st));
    H.o_bbeg = o_bbeg; H.o_bcnt = o_bcnt; H.o_runb0 = o_runb0; H.o_si = o_si; H.o_sj0 = o_sj0;
    H.o_sj1 = o_sj1; H.o_rj0 = o_rj0; H.o_rns = o_rns; H.o_rs0 = o_rs0; H.o_ord = o_ord;
    H.o_c0 = o_c0;
    H.run_seg0 = run_seg0;
    H.seg_len = S;
    H.seg_tj = span;
    H.sizes = E.run_n;
    H.nblocks = nb;
    H.nseg = ns;
    H.part_len = plen;
    H.bands = bands;
    return SPHX_OK;
}

static int sym_run(Engine &E, hipStream_t st) {
    SymHost &H = *E.grav.sym;
    int32_t *d = H.dev_i;
    int s0 = 0, s1 = H.nseg;
    int64_t p0 = 0, p1 = E.n;
    if (E.rrun >= 0) {
        s0 = H.run_seg0[E.rrun];
        s1 = H.run_seg0[E.rrun + 1];
        p0 = E.rp0;
        p1 = E.rp1;
    }
    if (s1 <= s0) return SPHX_OK;
    const bool blk = E.blk_t1 > 0;
    if (blk && H.act_cap < H.nblocks) {
        SPHX_CHECK(E.alloc((void **)&H.act, sizeof(int32_t) * (H.nblocks + 64)));
        H.act_cap = H.nblocks + 64;
    }
    k_tile_boxes<<<H.nblocks, 64, 0, st>>>(E.t, d + H.o_runb0, H.nblocks, SYM_B, E.s.m, H.bbox,
                                           blk ? E.s.bin : nullptr, E.blk_B, E.blk_t1, H.act);
    const unsigned pg = (unsigned)((p1 - p0 + 255) / 256);
    k_grav_sym_zero<<<pg, 256, 0, st>>>(p0, p1, E.s.gx, E.s.gy, E.s.gz);
    SPHX_HIP(hipGetLastError());
    const double *eps = E.p.eps_from_h ? E.s.h : E.s.eps;
    // the kernel-time phase spans the pair launches (and, with several bands,
    // the reductions between them)
    bool marked = false;
    int left = 0;
    for (const SymHost::Band &bd : H.bands) left += std::min(bd.seg_hi, s1) > std::max(bd.seg_lo, s0);
    for (const SymHost::Band &bd : H.bands) {
        const int a = std::max(bd.seg_lo, s0), b = std::min(bd.seg_hi, s1);
        if (b <= a) continue;
        const int64_t segs = bd.seg_hi - bd.seg_lo, ql = (int64_t)SYM_B * segs;
        double *q = H.q;
        float *part = H.budget >= 0 ? reinterpret_cast<float *>(H.pool + 3 * ql) : H.part;
        const int64_t pl = bd.slots;
        SymTabs T{d + H.o_bbeg, d + H.o_bcnt, d + H.o_runb0, d + H.o_si + a, d + H.o_sj0 + a,
                  d + H.o_sj1 + a, H.row_off, d + H.o_rj0, d + H.o_rns, d + H.o_rs0, H.bbox,
                  part, part + pl, part + 2 * pl, q, q + ql, q + 2 * ql, a, bd.seg_lo, bd.p_base,
                  blk ? H.act : nullptr, E.s.run, E.t.r_pbeg, E.n, d + H.o_ord, d + H.o_c0 + a,
                  H.seg_tj};
        if (!marked) SPHX_CHECK(E.tmark(st, PH_GRAVK, 0));
        marked = true;
        k_grav_sym<<<(unsigned)(b - a), 64, 0, st>>>(T, E.s.x, E.s.y, E.s.z, E.s.m, eps);
        SPHX_HIP(hipGetLastError());
        if (--left == 0) SPHX_CHECK(E.tmark(st, PH_GRAVK, 1));
        k_grav_sym_reduce<<<pg, 256, 0, st>>>(T, E.p.G, p0, p1, bd.row_lo, bd.row_hi, E.s.gx,
                                               E.s.gy, E.s.gz);
        SPHX_HIP(hipGetLastError());
    }
    return SPHX_OK;
}

int sphx_gravity_prepare(Engine &E, const GravConfig &cfg, hipStream_t st) {
    if (E.n == 0) return SPHX_OK;
    if (cfg.ipl == 8) return sym_prepare(E, cfg, st);
    if (E.n == 0) return SPHX_OK;
    GravHost &H = E.grav;
    int ib = 256 * cfg.ipl;
    bool rebuild = H.ipl != cfg.ipl || H.tj != cfg.tj || H.sizes != E.run_n;
    if (rebuild) {
        std::vector<int32_t> beg, cnt, run, tbeg;
        H.run_ib0.clear();
        int64_t off = 0;
        int tiles = 0;
        for (int r = 0; r < E.R; ++r) {
            int64_t nr = E.run_n[r];
            H.run_ib0.push_back((int32_t)beg.size());
            for (int64_t a = 0; a < nr; a += ib) {
                beg.push_back((int32_t)(off + a));
                cnt.push_back((int32_t)std::min<int64_t>(ib, nr - a));
                run.push_back(r);
            }
            tbeg.push_back(tiles);
            tiles += (int)((nr + cfg.tj - 1) / cfg.tj);
            off += nr;
        }
        H.run_ib0.push_back((int32_t)beg.size());
        int nb = (int)beg.size();
        int64_t need = 3 * (int64_t)nb + E.R;
        if (need > H.dev_cap) {
            SPHX_CHECK(E.alloc((void **)&H.dev, sizeof(int32_t) * (need + 1024)));
            H.dev_cap = need + 1024;
        }
        std::vector<int32_t> all;
        all.insert(all.end(), beg.begin(), beg.end());
        all.insert(all.end(), cnt.begin(), cnt.end());
        all.insert(all.end(), run.begin(), run.end());
        all.insert(all.end(), tbeg.begin(), tbeg.end());
        if (!all.empty())
            SPHX_HIP(hipMemcpyAsync(H.dev, all.data(), sizeof(int32_t) * all.size(),
                                    hipMemcpyHostToDevice, st));
        if (8 * (int64_t)tiles > H.tbox_cap) {
            SPHX_CHECK(E.alloc((void **)&H.tbox, sizeof(double) * (8 * (int64_t)tiles + 1024)));
            H.tbox_cap = 8 * (int64_t)tiles + 1024;
        }
        H.ipl = cfg.ipl;
        H.tj = cfg.tj;
        H.nblocks = nb;
        H.ntiles = tiles;
        H.sizes = E.run_n;
        SPHX_HIP(hipStreamSynchronize(st));
    }
    if (cfg.js > 1) {
        int64_t need = (int64_t)cfg.js * E.n * 4;
        if (need > H.part_cap) {
            SPHX_CHECK(E.alloc((void **)&H.part, sizeof(double) * need));
            H.part_cap = need;
        }
    }
    return SPHX_OK;
}

int sphx_gravity_run(Engine &E, const GravConfig &cfg_in, bool with_pot, hipStream_t st) {
    if (E.n == 0) return SPHX_OK;
    if (cfg_in.ipl == 8 && !with_pot) {
        SPHX_CHECK(sym_prepare(E, cfg_in, st));
        return sym_run(E, st);
    }
    // the symmetric sum carries no potential: that call uses the tiled sum
    const GravConfig cfg = cfg_in.ipl == 8 ? GravConfig{2, 512, 1} : cfg_in;
    SPHX_CHECK(sphx_gravity_prepare(E, cfg, st));
    GravHost &H = E.grav;
    int32_t *d = H.dev;
    // a run-restricted launch covers only that run's i-blocks (its j range is
    // its own run anyway)
    int ib0 = 0, ib1 = H.nblocks;
    if (E.rrun >= 0) {
        ib0 = H.run_ib0[E.rrun];
        ib1 = H.run_ib0[E.rrun + 1];
    }
    if (ib1 <= ib0) return SPHX_OK;
    GravTabs gt{d + ib0, d + H.nblocks + ib0, d + 2 * H.nblocks + ib0, E.t.r_pbeg, E.t.r_pend,
                d + 3 * H.nblocks, H.tbox, cfg.js};
    if (H.ntiles > 0) {
        k_tile_boxes<<<H.ntiles, 64, 0, st>>>(E.t, d + 3 * H.nblocks, H.ntiles, cfg.tj, E.s.m,
                                              H.tbox, nullptr, 0, 0, nullptr);
        SPHX_HIP(hipGetLastError());
    }
    const double *eps = E.p.eps_from_h ? E.s.h : E.s.eps;
    GravIO io{E.s.x, E.s.y, E.s.z, E.s.m, eps, E.p.G, E.s.gx, E.s.gy, E.s.gz, E.s.pot, H.part, E.n,
              E.s.bin, E.blk_t1 > 0 ? E.blk_B : 0, E.blk_t1};
    int nblk = ib1 - ib0;
    int rc;
    SPHX_CHECK(E.tmark(st, PH_GRAVK, 0));
    switch (cfg.ipl * 10000 + cfg.tj) {
        case 10256: rc = launch_grav<1, 256>(gt, io, nblk, with_pot, st); break;
        case 10512: rc = launch_grav<1, 512>(gt, io, nblk, with_pot, st); break;
        case 20256: rc = launch_grav<2, 256>(gt, io, nblk, with_pot, st); break;
        case 20512: rc = launch_grav<2, 512>(gt, io, nblk, with_pot, st); break;
        case 21024: rc = launch_grav<2, 1024>(gt, io, nblk, with_pot, st); break;
        case 40256: rc = launch_grav<4, 256>(gt, io, nblk, with_pot, st); break;
        case 40512: rc = launch_grav<4, 512>(gt, io, nblk, with_pot, st); break;
        default:
            sphx_set_error("unsupported gravity configuration");
            return SPHX_ERR_ARG;
    }
    SPHX_CHECK(rc);
    SPHX_CHECK(E.tmark(st, PH_GRAVK, 1));
    if (cfg.js > 1) {
        int64_t p0 = 0, p1 = E.n;
        if (E.rrun >= 0) {
            p0 = E.rp0;
            p1 = E.rp1;
        }
        k_grav_reduce<<<(unsigned)((p1 - p0 + 255) / 256), 256, 0, st>>>(
            H.part, cfg.js, E.n, p0, p1, E.s.gx, E.s.gy, E.s.gz, with_pot ? E.s.pot : nullptr);
        SPHX_HIP(hipGetLastError());
    }
    return SPHX_OK;
}

// Launch the pair-rate reference: blocks x 256 lanes, 16 pairs per lane per
// step; out holds blocks * 256 floats.
int sphx_pair_probe_launch(int64_t blocks, int iters, float *out, hipStream_t st) {
    k_pair_probe<<<(unsigned)blocks, 256, 0, st>>>(out, iters, 1e-3f);
    SPHX_HIP(hipGetLastError());
    return SPHX_OK;
}
