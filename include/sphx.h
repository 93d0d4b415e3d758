/*
 * sphx — C ABI of the MI355X (gfx950) SPH step library (libsphx.so).
 *
 * Drop-in boundary for the per-particle SPH step of hassanfv/SPH_2's
 * GPU_sph / Evrard_collapse drivers.  Reference paths below are relative to
 * 11_Dec_2022/ of that repository:
 *   LIBSX  = GPU_sph/IC/libsx2_2t.py            (FP64 Python/Numba library)
 *   LIBSX3 = GPU_sph/work_on_smoothing/libsx2_2t.py
 *   SHEAR  = GPU_sph/work_on_smoothing/shear.py
 *   PHOTOPY= GPU_sph/Cooling_v1/photolibs3.py
 *   LIBV2  = GPU_sph/Important_Test_Simulations/Evrard_collapse/myCppSPHLibs_v2.h
 *   DRV2   = GPU_sph/Important_Test_Simulations/Evrard_collapse/hfv_sph_gpu_v2_evrard.cu
 *
 * Conventions
 *   - Every array pointer is a DEVICE pointer (hipMalloc'd or a torch-ROCm
 *     tensor's data_ptr); positions and velocities are C-contiguous (n, 3)
 *     float64 like the reference's numpy arrays; per-particle scalars are (n,)
 *     float64; counts are int32.
 *   - `stream` is a hipStream_t (NULL = legacy default stream).  Calls are
 *     asynchronous unless stated; results are valid after the stream syncs.
 *   - Status: 0 ok; > 0 completed with capped iterations (see the counters);
 *     < 0 error (bad argument, out of memory, HIP failure).  Nothing is
 *     printed; sphx_last_error() returns a message for the calling thread.
 *   - Precision mode: SPHX_MIXED runs Morton-bucketed pair loops with FP32
 *     pair arithmetic and compensated FP32 accumulation, FP64 neighbour
 *     decisions, integrator and cooling solve.  SPHX_FP64 runs the plain
 *     baseline: one thread per particle, brute-force FP64 pair loops in the
 *     style of LIBV2, no LDS and no cells.
 *   - Rows [row_beg, row_end) select the target particles, as the reference's
 *     `*_mpi(nbeg, nend, ...)` variants do (e.g. LIBSX:58-95); outputs hold
 *     row_end - row_beg rows.
 */
#ifndef SPHX_H
#define SPHX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPHX_OK 0
#define SPHX_CAPPED 1
#define SPHX_REFUSED 2
#define SPHX_ERR_ARG (-1)
#define SPHX_ERR_NOMEM (-2)
#define SPHX_ERR_HIP (-3)
#define SPHX_ERR_STATE (-4)

#define SPHX_MIXED 0
#define SPHX_FP64 1

/* Physics constants of the step (augsphx6.5x.py:472-484; DRV2:283-300). */
typedef struct sphx_params {
    double gamma;        /* adiabatic index, 5/3 */
    double alpha, beta;  /* artificial-viscosity coefficients (1, 2) */
    double eta;          /* M83 viscosity softening, 0.1 */
    double G;            /* gravitational constant, 1 */
    int32_t nngb;        /* target neighbour count, 64 */
    int32_t ndown, nup;  /* accepted band, 59..69 (DRV2:283-285) */
    int32_t max_iter;    /* h-iteration cap, 100 (LIBV2:55) */
    int32_t av_form;     /* 0 Monaghan-Gingold 1983, 1 Gadget signal velocity */
    int32_t balsara;     /* 1: Balsara switch on Pi_ij (LIBV2:400-405) */
    int32_t eps_from_h;  /* 1: gravitational softening eps_i = h_i */
    int32_t h_predictor; /* 1: Gadget-2 eq. 31 start of the h iteration (LIBSX3:960) */
    int32_t h_mode;      /* 0: neighbour-count band iteration; 1: kNN h = half the distance
                            of rank knn_kth, as do_smoothingX (LIBSX:762-783) each step of
                            augsphx6.5.py:510 (mixed precision only); 2: the reference's
                            own rule, steps of 0.002 h_prev from 2 h_prev (or the eq. 31
                            start) until ndown <= k <= nup, the largest radius after
                            max_iter + 1 passes (LIBSX3:947-1006, LIBV2:7-64); 3: the same
                            with steps of 0.02 h_prev (do_smoothingX4); 4: the LIBV7
                            driver's rule, steps of 0.01 h_prev halved on every jump
                            across the band, the last radius after 501 passes, counts
                            stopping at nup + 1 (LIBV7:255-328); 5: h_smooth_fast,
                            h -+= 0.003 h from h_prev (no predictor) on the strict
                            count sqrt(r^2) < 2h, the last h after 10000 passes
                            (LIBSX:1133-1184) */
    int32_t knn_kth;     /* rank for h_mode 1 (64, self included at rank 0) */
    double av_w_eps;     /* Gadget AV: w = v_ij.r_ij / (r + av_w_eps); 1e-5 in LIBSX:158-169
                            and LIBV2, 1e-15 in augsphx6.5.py PI_ij */
    double h_min;        /* h_mode 5 only, 0 = none: after the eleventh step an h below
                            h_min becomes h_min and the iteration stops
                            (h_smooth_fast_h_minimum_set, LIBSX:1192-1245) */
} sphx_params;

/* Tabulated cooling: Lambda(T) on increasing T[ntab] (K) in erg cm^3 s^-1,
 * linear in T inside a bin; heating `heat` erg/s per hydrogen atom; code ->
 * cgs unit factors.  T and L are HOST pointers, copied by the callee. */
typedef struct sphx_cooling {
    double unit_u_cgs, unit_rho_cgs, unit_time_s;
    double XH, mu, gamma, heat, tol;
    int32_t max_iter, max_bracket, ntab, pad;
    const double *T;
    const double *L;
} sphx_cooling;

/* 2-D (u, rho) cooling table (the applyCooling method of
 * GPU_sph/With_cooling/Model_A/myCppSPHLibs_v3.h:794-859): DEVICE arrays of
 * strictly increasing cgs nodes and the cgs energy lost over dt_ref_s at each
 * node, du[j * nrho + i] for (u_nodes[j], rho_nodes[i]); code -> cgs factors. */
typedef struct sphx_cool_grid {
    int64_t nu, nrho;
    const double *u_nodes;
    const double *rho_nodes;
    double *du;
    double dt_ref_s;
    double unit_u_cgs, unit_rho_cgs, unit_time_s;
} sphx_cool_grid;

/* 5-D CHIMES table (the hcoolingx method of GPU_sph_Type_cooling_II/
 * hfvCLibs_v7.h:1077-1240, as its drivers hfvGpuSphTyp_v7.cu:35-79 and
 * hfvGpuSph_mgpu_v5.cu load it from coolHeatGridZ.bin): DEVICE float32 arrays.
 * Axes: kpc[n_kpc] (radius from the origin), nh[n_nh] (cm^-3), time_s[n_time]
 * (s), all strictly increasing; u_evol and mu are [n_kpc][n_t][n_nh][n_z]
 * [n_time], metalz [n_kpc][n_t][n_nh][n_z][n_m][n_time].  u_evol[.., 0] holds
 * the node energies (erg/g).  z_index is the metallicity node used (the
 * reference fixes 1), xh, m_h the hydrogen fraction and mass (0.70,
 * 1.6726e-24), u_max_cgs the energy cap (2.0287e18, T = 1e10 K), kpc_cm the
 * kpc in cm (3.086e21); code -> cgs factors for length, density, energy, time. */
typedef struct sphx_chimes_table {
    int32_t n_kpc, n_nh, n_z, n_t, n_m, n_time;
    int32_t z_index, pad;
    const float *kpc, *nh, *time_s;
    const float *u_evol, *mu, *metalz;
    double xh, m_h, u_max_cgs, kpc_cm;
    double unit_length_cm, unit_rho_cgs, unit_u_cgs, unit_time_s;
} sphx_chimes_table;

/* The analytic H/He cooling network of GPU_sph/Cooling_v1/photolibs3.py:
 * 140-543 (photoionisation equilibrium of H and He under a UV background,
 * Cen 1992 collisional rates, the T(u) fixed point, and DoCooling_h's
 * backward-Euler bracket x/÷1.1 and bisection), in FP64.  The background
 * J_nu = J21 1e-21 (nu_H0 / nu) erg s^-1 cm^-2 Hz^-1 sr^-1 (the reference's
 * Jcoeff: 1e-5 in photolibs3.py, 50 in the GPU drivers' myPhotoLibsGPU.h:
 * 166-296) is integrated over a frequency grid from each ionisation edge to
 * 6e16 Hz: n_grid 0 as photolibs3.py (np.arange with step 2e14 Hz), n_grid
 * > 0 that many points per edge (200 in myPhotoLibsGPU.h). */
typedef struct sphx_network {
    double XH;                                  /* hydrogen mass fraction, 0.76 */
    double J21;                                 /* background strength (Jcoeff) */
    double unit_u_cgs, unit_rho_cgs, unit_time_s;  /* code -> cgs (1 for cgs inputs) */
    double tol;                                 /* bisection ends at |du/u| <= tol, 1e-4 */
    int32_t n_grid;                             /* frequency grid, see above */
    int32_t max_iter;                           /* bisection cap, 100 (the reference's) */
    int32_t max_bracket;                        /* bracket-step cap (the reference has none) */
    int32_t pad;
} sphx_network;

typedef struct sphx_counters {
    int64_t n;              /* particles resident */
    int64_t steps;          /* steps taken since load */
    int64_t h_iters_max;    /* largest h-iteration count in the last step */
    int64_t h_iters_sum;    /* total h-iteration counts in the last step */
    int64_t h_capped;       /* particles whose h iteration did not converge */
    int64_t cool_capped;    /* particles whose cooling solve hit its cap */
    int64_t cand_buckets;   /* candidate bucket pairs in the last neighbour build */
    int64_t buckets;        /* number of 64-particle buckets */
    int64_t bh_overflow;    /* Barnes-Hut walk groups whose cell stack overflowed */
    int64_t nonfinite;      /* particles whose u, v or x was not finite after the last step */
    int64_t bin_floor;      /* block time steps: particles whose criterion asked for a step
                               shorter than the shortest level in the last substep */
    int64_t cool_iters;     /* bisection iterations of the implicit cooling solve, summed
                               over the particles of the last step */
    int64_t cool_bisected;  /* particles whose cooling solve bracketed and bisected a root in
                               the last step (the others left at the table floor or unchanged) */
    int64_t nbr_overflow;   /* buckets whose neighbour lists did not fit in the last step
                               (their div/curl and hydro sums came from the candidate walk) */
} sphx_counters;

typedef struct sphx_diag {
    double ke, ut, pe;      /* kinetic, thermal, potential energy */
    double px, py, pz;      /* linear momentum */
    double lx, ly, lz;      /* angular momentum about the origin */
    double mass;
} sphx_diag;

/* The positional params.GPU file of the reference's GPU drivers
 * (GPU_sph/Important_Test_Simulations/Evrard_collapse/
 * hfv_sph_gpu_v2_evrard.cu:38-60). */
typedef struct sphx_params_gpu {
    int64_t N;
    double c_0, gamma, Rcld_in_pc, Rcld_in_cm, Mcld_in_g, mu, Mach, grav_const_in_cgs, G;
} sphx_params_gpu;

typedef struct sphx_ctx sphx_ctx;
typedef struct sphx_plan sphx_plan;
typedef struct sphx_batch sphx_batch;

const char *sphx_last_error(void);
const char *sphx_version(void);
void sphx_default_params(sphx_params *p);

/* ---------------- stateless drop-ins (scratch owned by a context) ---------- */
int sphx_ctx_create(int device, sphx_ctx **out);
int sphx_ctx_destroy(sphx_ctx *ctx);

/* getDensity(pos, m, h) -> rho (LIBSX:17-51; LIBV2:69-114); nngb (nullable)
 * receives the neighbour count under r_ij <= h_i + h_j. */
int sphx_getDensity(sphx_ctx *ctx, int mode, int64_t n, const double *pos, const double *m,
                    const double *h, int64_t row_beg, int64_t row_end, double *rho,
                    int32_t *nngb, void *stream);

/* Adiabatic EOS: getPressure (LIBSX:100-105) and getCsound (LIBV2:783-789). */
int sphx_eos(sphx_ctx *ctx, int64_t n, const double *rho, const double *u, double gamma,
             double *P, double *c, void *stream);

/* smoothing_fast_new_mpi3(nbeg, nend, pos, h, Nngb_previous, divV, dt)
 * (LIBSX3:946-1010; LIBV7:255-330; LIBV2:7-64).  k_prev/divV nullable: without
 * them the iteration starts at 2*h_prev.  Outputs h and the neighbour count k
 * (self included, r_ij <= 2h_i). */
int sphx_smoothing(sphx_ctx *ctx, int mode, int64_t n, const double *pos, const double *h_prev,
                   const int32_t *k_prev, const double *divV, double dt, const sphx_params *p,
                   int64_t row_beg, int64_t row_end, double *h_out, int32_t *k_out,
                   int32_t *iters_out, void *stream);

/* div_curlVel(pos, v, rho, m, h) -> (divV signed, |curlV|) (SHEAR:8-69; LIBV2:247-322). */
int sphx_div_curlVel(sphx_ctx *ctx, int mode, int64_t n, const double *pos, const double *v,
                     const double *rho, const double *m, const double *h, int64_t row_beg,
                     int64_t row_end, double *divV, double *curlV, void *stream);

/* getAcc_sph and get_dU in one pair pass (LIBSX:110-186 and 363-430;
 * SHEAR:151-235 and 334-406 with Balsara; LIBV2:327-417 and 634-719).
 * divV/curlV are read only when p->balsara; acc (rows x 3) and dudt nullable. */
int sphx_getAcc_sph(sphx_ctx *ctx, int mode, int64_t n, const double *pos, const double *v,
                    const double *rho, const double *P, const double *c, const double *h,
                    const double *m, const double *divV, const double *curlV,
                    const sphx_params *p, int64_t row_beg, int64_t row_end, double *acc,
                    double *dudt, void *stream);

/* getAcc_g_smth(pos, mass, G, epsilon) (LIBSX:633-671; LIBV2:119-168);
 * pot (nullable) receives phi_i = -G sum_j m_j phi_ij (the pair term of getPE,
 * LIBSX:593-629). */
int sphx_getAcc_g_smth(sphx_ctx *ctx, int mode, int64_t n, const double *pos, const double *m,
                       const double *eps, double G, int64_t row_beg, int64_t row_end,
                       double *acc, double *pot, void *stream);

/* Issue-rate ceiling of the direct sum's far pairs on this device: a kernel
 * with the far-pair instruction mix of the symmetric sum (3 sub, 1 mul + 2 fma
 * for r^2, rsq, 2 mul, 3 + 3 fma to both particles' sums), 8 i and 2 j per
 * lane in registers, no memory traffic, eight waves per SIMD.  Writes the pair
 * evaluations per second it sustains.  A measurement aid with no counterpart
 * in the reference: it is what the direct sum's pair rate is compared with.
 * iters in 1..65536 (16 pairs per lane per iteration); the fastest of three
 * timed launches after a warm-up.  Synchronous. */
int sphx_pair_issue_rate(sphx_ctx *ctx, int32_t iters, double *pairs_per_s, void *stream);

/* Barnes-Hut gravity with the spline softening of getAcc_g_smth: Morton-built
 * octree (cells = key ranges), monopole moments, acceptance when a cell's
 * distance from the 64-particle Morton group's box exceeds s/theta + |com -
 * centre|.  Replaces the octree of Particles_in_BH_Tree_II/Evrard_test/
 * bh_tree_iteration_v2.h:29-588.  stats (nullable, 8 int64): cells, depth,
 * pair interactions, stack overflows, largest interaction-list fill, device
 * microseconds of the tree build and of the walk, walk groups.
 * Synchronous; SPHX_ERR_STATE if a group's frontier overflowed. */
int sphx_bh_gravity(sphx_ctx *ctx, int64_t n, const double *pos, const double *m,
                    const double *eps, double G, double theta, int64_t row_beg, int64_t row_end,
                    double *acc, double *pot, int64_t *stats, void *stream);

/* Neighbour lists under r_ij <= h_i + h_j for rows [row_beg,row_end), ascending
 * j per row (the gate of LIBV2:89).  offsets (rows+1, int64) always written;
 * idx (int32) written when capacity suffices; *total receives the pair count.
 * Synchronous. */
int sphx_neighbors(sphx_ctx *ctx, int64_t n, const double *pos, const double *h,
                   int64_t row_beg, int64_t row_end, int64_t *offsets, int32_t *idx,
                   int64_t capacity, int64_t *total, void *stream);

/* Implicit per-particle cooling solve over the tabulated Lambda(T)
 * (DoCooling_h, PHOTOPY:489-543).  iters nullable.  Returns SPHX_CAPPED when
 * any particle hit max_iter. Synchronous. */
int sphx_cool(sphx_ctx *ctx, int64_t n, const double *rho, const double *u, double dt,
              const sphx_cooling *cool, double *u_out, int32_t *iters, void *stream);

/* The six rates of the network's background (RadiationField,
 * photolibs3.py:230-329): photoionisation rates of H0, He0, He+ (s^-1) and
 * their heating rates (erg s^-1), into out[6].  Host only. */
int sphx_network_rates(const sphx_network *net, double *out);
/* DoCooling_h (photolibs3.py:489-543) on device arrays in the code units of
 * net: u_out = the backward-Euler u after dt, with the net rate of
 * coolingRateFromU_h (:470-486, T(u) clamped to 1e4..1e8 K).  iters
 * (nullable) receives the bisection iterations.  u_out may alias u.  Returns
 * SPHX_CAPPED when a solve hit max_iter or max_bracket.  Synchronous. */
int sphx_cool_network(sphx_ctx *ctx, int64_t n, const double *rho, const double *u, double dt,
                      const sphx_network *net, double *u_out, int32_t *iters, void *stream);

/* applyCooling (LIBV3 myCppSPHLibs_v3.h:794-859) on device arrays: u, rho to
 * cgs, clamped inside the table, bin = first j with g[j] <= x <= g[j+1] (the
 * reference's linear scans, found here by binary search), u_new = u_nodes[j]
 * - (dt_s/dt_ref_s) du[j, i], floored at u_nodes[0] (the reference has no
 * floor).  u_out may alias u_in. */
int sphx_cool_grid_apply(sphx_ctx *ctx, int64_t n, const double *rho, const double *u_in,
                         double dt, const sphx_cool_grid *grid, double *u_out, void *stream);
/* Fill grid->du from the implicit solve of sphx_cool over dt_ref_s at every
 * node (the role of GPU_sph/Cooling_v1/createCoolingGridX2.cu:9-35). */
int sphx_cool_grid_build(sphx_ctx *ctx, const sphx_cooling *cool, const sphx_cool_grid *grid,
                         void *stream);

/* hcoolingx (hfvCLibs_v7.h:1077-1240) on device arrays, in FP64 on the float32
 * table: nearest kpc and n_H nodes (first node >= value, one down when the
 * lower node is strictly closer, past the end the last), first T node whose
 * energy (time index 0) is >= u and first time node >= dt (else the last),
 * energy interpolated linearly in u between T nodes, capped at u_max_cgs.
 * ion_out (n x n_m float32) and mu_out receive the species fractions and mu
 * at the chosen nodes (both nullable; the reference kernel writes the
 * fractions, its CPU twin hcooler in cooling_grids_use_EC2/checker1.py:53-146
 * returns mu).  pos is (n, 3).  u_out may alias u_in. */
int sphx_chimes_apply(sphx_ctx *ctx, int64_t n, const double *pos, const double *rho,
                      const double *u_in, double dt, const sphx_chimes_table *table,
                      double *u_out, float *ion_out, double *mu_out, void *stream);

/* getKE (LIBSX:579-588), thermal energy sum m u, momentum and angular
 * momentum; pe only when pos/eps given (getPE, LIBSX:593-629).  Synchronous. */
int sphx_diagnostics(sphx_ctx *ctx, int64_t n, const double *pos, const double *v,
                     const double *m, const double *u, const double *eps, double G,
                     sphx_diag *out, void *stream);

/* do_smoothingX((pos, pos[row_beg:row_end])) (LIBSX:762-783; AUG:305-326):
 * h = 0.5 x the kth smallest distance (0-based, self included; the reference
 * uses kth = 64) from each row to all n particles, exact in FP64 (bit-equal
 * to sorting the distances).  h_guess (device, n) seeds the search radius
 * only.  Returns SPHX_CAPPED with NaN rows if a row's search did not resolve. */
int sphx_knn_h(sphx_ctx *ctx, int64_t n, const double *pos, const double *h_guess, int32_t kth,
               int64_t row_beg, int64_t row_end, double *h_out, void *stream);

/* Global adaptive time step: dt_array (LIBV2:491-522) plus the driver's host
 * reduction (DRV2:517-561), done on the device without the O(N^2) v_sig loop
 * or the six host copies.  acc = total, acc_g = gravitational acceleration
 * (n, 3); divV signed or absolute.  out (device, 5 doubles) = dt, dt_f,
 * dt_kin, dt_cfl, dt_dens with dt = 0.25 min(...) clamped to [dt_min, dt_max]
 * (the driver: C_CFL = 0.25, [1e-5, 1e-3]). */
int sphx_global_dt(sphx_ctx *ctx, int64_t n, const double *h, const double *c,
                   const double *acc, const double *acc_g, const double *divV, double c_cfl,
                   double dt_min, double dt_max, double *out, void *stream);

/* sphx_global_dt without a context and without any allocation, for callers
 * that cannot allocate device memory (code under hipGraph capture, serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * Same kernels and result as sphx_global_dt; their per-bucket extrema go to a
 * fixed 5 MiB device array of the library (one per device, held until the
 * process ends), so n is limited to 8,388,608 particles (SPHX_ERR_ARG beyond).
 * `device` is made current for the call.  Calls that overlap on one device
 * share that array and race: serialise them (one stream, or events). */
int sphx_global_dt_noalloc(int device, int64_t n, const double *h, const double *c,
                           const double *acc, const double *acc_g, const double *divV,
                           double c_cfl, double dt_min, double dt_max, double *out, void *stream);

/* Morton keys (21 bits per axis, 63-bit key) over the bounding box of pos,
 * then a stable LSD radix sort; keys_out[i] and perm_out[i] = original index of
 * the i-th smallest key.  Synchronous. */
int sphx_morton_sort(sphx_ctx *ctx, int64_t n, const double *pos, uint64_t *keys_out,
                     int32_t *perm_out, void *stream);

/* ---------------- resident plan: the whole step on device ------------------ */
/* State stays on the device in Morton order between steps.  The step order is
 * DRV2:407-482: half kick, drift, h iteration, density + EOS, gravity,
 * div/curl, hydro acceleration + du/dt (same velocity), half kick, u
 * trapezoid, optional implicit cooling. */
int sphx_plan_create(int device, int mode, const sphx_params *p, const sphx_cooling *cool,
                     int64_t capacity, sphx_plan **out);
/* pos, vel (n,3), u, m, h_guess (n,), eps (nullable; used when !eps_from_h). */
int sphx_plan_load(sphx_plan *plan, int64_t n, const double *pos, const double *vel,
                   const double *u, const double *m, const double *h, const double *eps,
                   void *stream);
/* sphx_plan_load with the particle buffers left in the caller's ownership
 * (the driver's own device arrays, DRV2:150-189): pos, vel (n,3), u and h (n,)
 * become the run's state between calls.  Every sphx_plan_init /
 * sphx_plan_step reads them at its start (edits made between calls take
 * effect) and writes the state back into them at its end, in the caller's
 * particle order; m and eps are read once here.  The plan keeps its own
 * Morton-ordered working copy, because the sort reorders the state every step:
 * binding costs a gather and a scatter of 8 doubles per particle per call
 * (about 0.1 ms at N = 1,048,576), not per step.  sphx_plan_load or
 * sphx_plan_restore ends the binding. */
int sphx_plan_bind(sphx_plan *plan, int64_t n, double *pos, double *vel, double *u,
                   const double *m, double *h, const double *eps, void *stream);
/* Forces at t = 0 (DRV2:311-392 without its Euler u update). */
int sphx_plan_init(sphx_plan *plan, void *stream);
/* n_steps KDK steps at fixed dt.  use_cooling: 0 none, 1 the implicit solve
 * on the plan's Lambda(T) table, 2 the 2-D (u, rho) table of
 * sphx_plan_set_cooling_grid, 3 the 5-D CHIMES table of
 * sphx_plan_set_chimes_table, 4 the analytic H/He network of
 * sphx_plan_set_cooling_network. */
int sphx_plan_step(sphx_plan *plan, int32_t n_steps, double dt, int32_t use_cooling,
                   void *stream);
/* Copy state out in the original particle order; any pointer may be NULL.
 * acc = total acceleration (n,3); acc_g = gravity part (n,3). */
int sphx_plan_store(sphx_plan *plan, double *pos, double *vel, double *u, double *h,
                    double *rho, double *acc, double *acc_g, double *dudt, int32_t *k,
                    void *stream);
int sphx_plan_query(sphx_plan *plan, sphx_counters *out);
/* Energies and momenta of the resident state (pe needs the potential pass). */
int sphx_plan_diagnostics(sphx_plan *plan, int32_t with_pe, sphx_diag *out, void *stream);
int sphx_plan_destroy(sphx_plan *plan);

/* Gravity launch shape of a plan or context: ipl particles per lane (1, 2, 4),
 * j-tile tj (256, 512, 1024), j split js >= 1 (more workgroups for small N);
 * or ipl = 8, the symmetric sum: (8, 512, S) with S block pairs per wave, or
 * (8, 256, 1) / (8, 128, 1) with each block pair split into parts of tj j.
 * No reference counterpart: the reference launches acc_g with a fixed
 * <<<ceil(N/256), 256>>> (DRV2:280-281); the autotuner sets this. */
int sphx_plan_set_gravity_config(sphx_plan *plan, int32_t ipl, int32_t tj, int32_t js);
int sphx_ctx_set_gravity_config(sphx_ctx *ctx, int32_t ipl, int32_t tj, int32_t js);

/* Per-phase HIP-event timing of sphx_plan_step (the reference times its
 * passes with std::chrono around synchronised launches,
 * GPU_sph_Type/Evrard_collapse/individual_dt_not_completed/AWS/main.cu:548-602).
 * Phases: 0 kick+drift+sort, 1 h iteration, 2 density+EOS, 3 gravity,
 * 4 div/curl, 5 hydro+du/dt, 6 kick+u+cooling.  set_timing resets the sums;
 * timing returns summed milliseconds and step counts per phase. */
int sphx_plan_set_timing(sphx_plan *plan, int32_t on);
int sphx_plan_timing(sphx_plan *plan, double *ms, int64_t *count, int32_t nphase);

/* Block time steps (levels > 0; 0 switches them off): particle i steps with
 * dt / 2^b_i, b_i in 0..levels, where dt is sphx_plan_step's argument (the
 * longest step, fixed until the next sphx_plan_init) and each sphx_plan_step
 * step is one substep dt / 2^levels.  b_i is the longest level whose step is
 * at most the per-particle form of the global criterion (0.25 min(sqrt(h/|a_g|),
 * sqrt(h/|a|), c_cfl h/c, c_cfl h/|h div v / 3|)); a particle moves to a
 * shorter level at once, to a longer one a level at a time where the substep
 * count allows.  Kick-drift-kick per particle: every particle drifts every
 * substep, the SPH passes run for all, gravity only for the particle blocks
 * holding a particle whose step ends.  The reference's prototype of individual
 * steps (GPU_sph_Type/Evrard_collapse/individual_dt_not_completed/AWS/
 * main.cu:472-520 with hfvCppLibs_indiv_dt.h:819-980) is unfinished; this is
 * this library's own scheme.  Mixed precision; not with the adaptive global
 * dt, graphs or checkpoints.  Call before the first step after init. */
int sphx_plan_set_block_steps(sphx_plan *plan, int32_t levels, double c_cfl);
/* Level b_i of every particle (int32, n, original order; device array). */
int sphx_plan_block_bins(sphx_plan *plan, int32_t *out, void *stream);

/* Gravity of the step from the Barnes-Hut tree of sphx_bh_gravity instead of
 * the direct sum (theta > 0; 0 restores the direct sum).  Mixed precision,
 * single-run plans; replaces the BH gravity call of the reference's tree
 * drivers (Particles_in_BH_Tree_II/Evrard_test/bh_tree_iteration_v2.h:29-588
 * with theta = 0.5).  Steps with the tree are not replayed from graphs. */
int sphx_plan_set_gravity_bh(sphx_plan *plan, double theta);

/* Attach a 2-D cooling table (copied); runs stepped with use_cooling = 2
 * cool by table lookup instead of the implicit solve (DRV3: hfv_sph_gpu_v3.cu:
 * 496-508, 600-609 apply it after the u update). */
int sphx_plan_set_cooling_grid(sphx_plan *plan, const sphx_cool_grid *grid);

/* Attach a 5-D CHIMES table (copied); runs stepped with use_cooling = 3 take
 * their energy from the table after the u update, as hfvGpuSphTyp_v7.cu:
 * 665-673 calls hcoolingx after u_updater.  The plan keeps each particle's
 * species fractions of its last lookup; sphx_plan_chimes_ions copies them out
 * (n x n_m float32, original particle order; zero before the first lookup). */
int sphx_plan_set_chimes_table(sphx_plan *plan, const sphx_chimes_table *table);
/* Attach the analytic H/He network (copied, rates evaluated here); runs
 * stepped with use_cooling = 4 solve DoCooling_h after the u update, as the
 * GPU drivers apply their cooling after the u update
 * (With_cooling/Model_A/hfv_sph_gpu_v3.cu:600-609). */
int sphx_plan_set_cooling_network(sphx_plan *plan, const sphx_network *net);
int sphx_plan_chimes_ions(sphx_plan *plan, float *out, void *stream);

/* Shift (-3..3) of the Morton level at which 64-particle buckets are cut into
 * segments for the neighbour passes (default 0: level floor(log8(n/64)) - 1).
 * A tuning knob only: every sum is met in the same order at any level, so
 * results do not change.  No counterpart in the reference (brute-force pair
 * loops, LIBV2:27-38). */
int sphx_plan_set_cell_level(sphx_plan *plan, int32_t bias);
/* Neighbour lists (on by default): the density pass records each particle's
 * accepted neighbours in the order its candidate walk meets them, and the
 * div/curl and hydro passes iterate those lists instead of walking the
 * candidates again.  The same pairs in the same order: results are
 * bit-identical either way; off runs the walks (a tuning and test knob).  A
 * particle with more than 128 such neighbours sends its bucket back to the
 * walk (counted in nbr_overflow).  No counterpart in the reference, whose
 * kernels test every pair (LIBV2:247-417). */
int sphx_plan_set_neighbor_lists(sphx_plan *plan, int32_t on);
int sphx_batch_set_cell_level(sphx_batch *b, int32_t bias);

/* Adaptive global time step inside the step loop (DRV2:517-561): after each
 * step's forces the device evaluates the criterion of sphx_global_dt and
 * uses it for the next step, with no host round trip.  The dt argument of
 * the first sphx_plan_step after enabling seeds the first step (the driver
 * starts from dt = 2e-4, DRV2:303); later dt arguments are ignored.
 * sphx_plan_time returns (host scalars) the simulated time -- the sum of the
 * dt each step used --, the next step's dt and (nullable, 4 doubles) the
 * last dt_f, dt_kin, dt_cfl, dt_dens.  Synchronous. */
int sphx_plan_set_adaptive_dt(sphx_plan *plan, int32_t on, double c_cfl, double dt_min,
                              double dt_max);
int sphx_plan_time(sphx_plan *plan, double *t, double *dt, double *dt_parts, void *stream);

/* Replay each step from a captured hipGraph (one graph per pointer parity of
 * the sorted state; recaptured when the run set changes).  Replaces the
 * launch-and-synchronise sequence of DRV2:407-482.  Timing mode disables it. */
int sphx_plan_set_graph(sphx_plan *plan, int32_t on);

/* ---------------- served runs: many simulations packed in one engine ------- */
/* No reference counterpart: the reference runs one simulation per process.
 * Each admitted run is an independent simulation (its own dt, cooling flag,
 * Morton order, buckets, gravity blocks) stored as a contiguous segment of
 * packed arrays; sphx_batch_step advances every live run with one launch per
 * pass.  A run's results do not depend on which other runs share the batch. */
int sphx_batch_create(int device, int mode, const sphx_params *p, const sphx_cooling *cool,
                      int64_t capacity, int32_t max_runs, sphx_batch **out);
/* Returns SPHX_REFUSED (2) without side effects when the particle capacity or
 * the run slots are exhausted.  Bootstraps the new run (forces at t = 0).
 * use_cooling as in sphx_plan_step (0 none, 1 solve, 2 2-D table, 3 CHIMES
 * table). */
int sphx_batch_admit(sphx_batch *b, int64_t n, const double *pos, const double *vel,
                     const double *u, const double *m, const double *h, const double *eps,
                     double dt, int32_t use_cooling, int64_t *run_id, void *stream);
/* Copies the run out in its caller order (any pointer may be NULL) and frees
 * its slot, compacting the packed arrays. */
int sphx_batch_retire(sphx_batch *b, int64_t run_id, double *pos, double *vel, double *u,
                      double *h, double *rho, double *acc, int32_t *k, void *stream);
/* Copies a live run's state out in its caller order without retiring it. */
int sphx_batch_store(sphx_batch *b, int64_t run_id, double *pos, double *vel, double *u,
                     double *h, double *rho, double *acc, int32_t *k, void *stream);
int sphx_batch_step(sphx_batch *b, int32_t n_steps, void *stream);
int sphx_batch_runs(sphx_batch *b, int64_t *ids, int64_t *sizes, int64_t *steps,
                    int32_t capacity, int32_t *count);
/* Per-run adaptive time step (as sphx_plan_set_adaptive_dt; the admit dt
 * seeds the first step) and the run's simulated time / next dt.  No
 * counterpart in the reference, whose driver runs one simulation. */
/* As sphx_plan_set_cooling_grid, for runs admitted with use_cooling = 2. */
int sphx_batch_set_cooling_grid(sphx_batch *b, const sphx_cool_grid *grid);
/* Gravity of the batch's runs from the symmetric direct sum (each pair
 * evaluated once for both particles; block_pairs = 512-particle block pairs
 * per wavefront) instead of the fixed tiled shape.  Its partial sums total
 * sphx_gravity_symmetric_bytes(n, block_pairs) per run of n particles (about
 * 12 n^2 / 1024 bytes); budget_bytes of device memory are reserved here and
 * the triangle of 512-particle block pairs is summed in bands of rows whose
 * partial sums fit them, each band's folded into FP64 accumulators before the
 * next (the sums do not depend on the banding).  A run whose largest row
 * does not fit, sphx_gravity_symmetric_min_bytes(n, block_pairs), is refused
 * (SPHX_REFUSED).  Only before the first admission.  A run's results do not
 * depend on the other runs either way.  No counterpart in the reference (one
 * simulation per process, acc_g of LIBV2:119-168 over all pairs).  Plans and
 * stateless calls band the same way with 16 GiB per band. */
int sphx_batch_set_gravity_symmetric(sphx_batch *b, int32_t block_pairs, int64_t budget_bytes);
int64_t sphx_gravity_symmetric_bytes(int64_t n, int32_t block_pairs);
int64_t sphx_gravity_symmetric_min_bytes(int64_t n, int32_t block_pairs);
/* As sphx_plan_set_chimes_table, for runs admitted with use_cooling = 3 (the
 * energy lookup only: batches keep no species fractions). */
int sphx_batch_set_chimes_table(sphx_batch *b, const sphx_chimes_table *table);
/* As sphx_plan_set_cooling_network, for runs admitted with use_cooling = 4. */
int sphx_batch_set_cooling_network(sphx_batch *b, const sphx_network *net);
int sphx_batch_set_adaptive_dt(sphx_batch *b, int64_t run_id, int32_t on, double c_cfl,
                               double dt_min, double dt_max, void *stream);
int sphx_batch_time(sphx_batch *b, int64_t run_id, double *t, double *dt, double *dt_parts,
                    void *stream);
int sphx_batch_set_graph(sphx_batch *b, int32_t on);
int sphx_batch_query(sphx_batch *b, sphx_counters *out);
/* Device bytes held per particle of capacity (the admission budget unit). */
int64_t sphx_batch_bytes_per_particle(void);
int sphx_batch_destroy(sphx_batch *b);

/* ---------------- files (HOST arrays; nothing here touches a GPU) ---------- *
 * The reference's formats, as its drivers and IC scripts read and write them
 * (CSV values here carry 17 significant digits instead of 6, which the
 * reference's stof still reads).  pos/vel are (n, 3) float64 host arrays. */

/* params.GPU (hfv_sph_gpu_v2_evrard.cu:38-60; IC/3.py:66-73). */
int sphx_read_params_gpu(const char *path, sphx_params_gpu *out);
int sphx_write_params_gpu(const char *path, const sphx_params_gpu *p);

/* IC / restart CSV rows x,y,z,vx,vy,vz,m,h,eps,u[,dudt]
 * (hfv_sph_gpu_v2_evrard.cu:75-131; Check_Restart_version/hfv_sph_gpu_v5.cu:
 * 348-396 read, 745-753 written).  dudt nullable (needs 11 columns);
 * rows_read (nullable) = data rows in the file. */
int sphx_read_ic_csv(const char *path, int64_t n, double *pos, double *vel, double *m, double *h,
                     double *eps, double *u, double *dudt, int64_t *rows_read);
int sphx_write_restart_csv(const char *path, int64_t n, const double *pos, const double *vel,
                           const double *m, const double *h, const double *eps, const double *u,
                           const double *dudt);
/* Snapshot CSV "x,y,z,vx,vy,vz,h,rho,u" (hfv_sph_gpu_v2_evrard.cu:495-514). */
int sphx_write_snapshot_csv(const char *path, int64_t n, const double *pos, const double *vel,
                            const double *h, const double *rho, const double *u);

/* Binary IC: int32 Typ[n], float32 x,y,z,vx,vy,vz,m,h,eps,u (IC_evrard_binary.py:
 * 60-82; read by GPU_sph_Type/Evrard_collapse/hfvCppLibs.h:8-63).  typ nullable
 * (written as 0). */
int sphx_write_ic_bin(const char *path, int64_t n, const int32_t *typ, const double *pos,
                      const double *vel, const double *m, const double *h, const double *eps,
                      const double *u);
int sphx_read_ic_bin(const char *path, int64_t n, int32_t *typ, double *pos, double *vel,
                     double *m, double *h, double *eps, double *u);
/* Binary snapshot: int32 N, int32 Typ[N], float32 x,y,z,vx,vy,vz,rho,h,u,mass
 * (hfvCppLibs.h:68-96).  Read with n_cap = 0 first to learn N. */
int sphx_write_snapshot_bin(const char *path, int64_t n, const int32_t *typ, const double *pos,
                            const double *vel, const double *rho, const double *h,
                            const double *u, const double *m);
int sphx_read_snapshot_bin(const char *path, int64_t n_cap, int64_t *n_out, int32_t *typ,
                           double *pos, double *vel, double *rho, double *h, double *u,
                           double *m);

/* coolHeatGridZ.bin, the CHIMES table file (written by GPU_sph_Type_cooling_II/
 * cooling_grids_use_EC2/step_2x.py:101-140, read by hfvGpuSphTyp_v7.cu:35-79):
 * int32 n_kpc, n_nh, n_z, n_t, n_m, n_time; float32 kpc, nh, z, temperatures,
 * time_s; float32 u_evol, mu, metalz in the layouts of sphx_chimes_table.
 * dims receives the six sizes in that order; with arrays null only dims is
 * read.  HOST arrays sized from dims. */
int sphx_read_chimes(const char *path, int32_t dims[6], float *kpc, float *nh, float *z,
                     float *temp, float *time_s, float *u_evol, float *mu, float *metalz);
int sphx_write_chimes(const char *path, const int32_t dims[6], const float *kpc,
                      const float *nh, const float *z, const float *temp, const float *time_s,
                      const float *u_evol, const float *mu, const float *metalz);

/* Exact checkpoints (no counterpart in the reference, whose restart CSV makes
 * the driver recompute forces from rounded values): every per-particle array
 * of a resident run in float64, in its internal Morton order, with its
 * parameters, step count, simulated time and dt.  A restored run continues
 * bit-identically to one that was never stopped.  The cooling table is not
 * stored: restore into a plan/batch created with the same table.  Restoring
 * into a batch needs the batch's sphx_params to equal the run's; it returns
 * SPHX_REFUSED when the capacity or run slots are exhausted. */
int sphx_plan_save(sphx_plan *plan, const char *path, void *stream);
int sphx_plan_restore(sphx_plan *plan, const char *path, void *stream);
int sphx_batch_save_run(sphx_batch *b, int64_t run_id, const char *path, void *stream);
int sphx_batch_restore_run(sphx_batch *b, const char *path, int64_t *run_id, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* SPHX_H */
