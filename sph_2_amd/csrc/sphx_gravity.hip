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
};

__global__ void k_tile_boxes(Tables t, const int32_t *__restrict__ r_tbeg, int ntiles_total,
                             int tj, double *__restrict__ tbox) {
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
    if (lane == 0) {
        for (int a = 0; a < 7; ++a) tbox[8 * tile + a] = v[a];
        tbox[8 * tile + 7] = 0.0;
    }
}

template <int IPL>
struct IState {
    int idx[IPL];  // global index of each i (-1 when the lane holds none)
    float x[IPL], y[IPL], z[IPL], e[IPL];
    float ax[IPL], ay[IPL], az[IPL], ph[IPL];
    KahanF kx[IPL], ky[IPL], kz[IPL], kp[IPL];
};

template <int IPL, bool POT>
__device__ __forceinline__ void far_pair(IState<IPL> &S, float4 j) {
#pragma unroll
    for (int p = 0; p < IPL; ++p) {
        float dx = j.x - S.x[p];
        float dy = j.y - S.y[p];
        float dz = j.z - S.z[p];
        float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        float ri = __builtin_amdgcn_rsqf(r2);
        float mri = j.w * ri;
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
        float e = 0.5f * (S.e[p] + ej);
        float f, ph;
        if (r2 >= 4.0f * e * e) {
            float ri = __builtin_amdgcn_rsqf(r2);
            f = ri * ri * ri;
            ph = ri;
        } else {
            float r = r2 > 0.0f ? __builtin_amdgcn_sqrtf(r2) : 0.0f;
            float ie = __builtin_amdgcn_rcpf(e);
            float q = r * ie;
            if (q <= 1.0f) {
                float q2 = q * q;
                f = ie * ie * ie * (4.0f / 3.0f - 1.2f * q2 + 0.5f * q2 * q);
                ph = ie * (1.4f - 2.0f * q2 * (1.0f / 3.0f - 0.15f * q2 + 0.05f * q2 * q));
            } else {
                float w = 2.0f - q;
                float w2 = w * w;
                float ri = __builtin_amdgcn_rcpf(r);
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

template <int IPL>
__device__ __forceinline__ void flush(IState<IPL> &S) {
#pragma unroll
    for (int p = 0; p < IPL; ++p) {
        S.kx[p].add(S.ax[p]); S.ax[p] = 0.f;
        S.ky[p].add(S.ay[p]); S.ay[p] = 0.f;
        S.kz[p].add(S.az[p]); S.az[p] = 0.f;
        S.kp[p].add(S.ph[p]); S.ph[p] = 0.f;
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
        if (far) {
#pragma unroll 8
            for (int q = 0; q < TJ; ++q) far_pair<IPL, POT>(S, jt[q]);
        } else {
            for (int q = 0; q < TJ; ++q) near_pair<IPL, POT>(S, jt[q], je[q], j0 + q);
        }
        flush<IPL>(S);
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

}  // namespace

int sphx_gravity_prepare(Engine &E, const GravConfig &cfg, hipStream_t st) {
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

int sphx_gravity_run(Engine &E, const GravConfig &cfg, bool with_pot, hipStream_t st) {
    if (E.n == 0) return SPHX_OK;
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
        k_tile_boxes<<<H.ntiles, 64, 0, st>>>(E.t, d + 3 * H.nblocks, H.ntiles, cfg.tj, H.tbox);
        SPHX_HIP(hipGetLastError());
    }
    const double *eps = E.p.eps_from_h ? E.s.h : E.s.eps;
    GravIO io{E.s.x, E.s.y, E.s.z, E.s.m, eps, E.p.G, E.s.gx, E.s.gy, E.s.gz, E.s.pot, H.part, E.n};
    int nblk = ib1 - ib0;
    int rc;
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
