// Host-side file formats: the reference's IC / snapshot / restart / parameter
// files, and the library's own exact checkpoint of a resident run.
//
// Reference formats (paths relative to 11_Dec_2022/ of hassanfv/SPH_2):
//   * params.GPU: ten whitespace-separated numbers read positionally --
//     N, c_0, gamma, Rcld_in_pc, Rcld_in_cm, Mcld_in_g, mu, Mach,
//     grav_const_in_cgs, G (GPU_sph/Important_Test_Simulations/
//     Evrard_collapse/hfv_sph_gpu_v2_evrard.cu:38-60; written by IC/3.py:66-73).
//   * IC CSV: no header, rows x,y,z,vx,vy,vz,m,h,eps,u (same driver:75-131;
//     written by IC/3.py:51-60); the restart CSV adds an 11th column dudt
//     (GPU_sph/With_metal_cooling/Check_Restart_version/hfv_sph_gpu_v5.cu:
//     745-753 written, 348-396 read).
//   * snapshot CSV: header "x,y,z,vx,vy,vz,h,rho,u" then rows
//     (hfv_sph_gpu_v2_evrard.cu:495-514).
//   * binary IC: int32 Typ[N] then float32 x,y,z,vx,vy,vz,m,h,eps,u [N] each,
//     no count (GPU_sph_Type/Evrard_collapse/IC_evrard_binary.py:60-82, read
//     by hfvCppLibs.h:8-63).
//   * binary snapshot: int32 N, int32 Typ[N], float32 x,y,z,vx,vy,vz,rho,h,u,
//     mass [N] each (hfvCppLibs.h:68-96).
// The reference writes CSV with the default 6 significant digits; here CSV
// values are written with 17 significant digits (still read by the
// reference's stof), so a CSV round trip is exact in FP64.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sphx_internal.h"

namespace {

struct File {
    FILE *f = nullptr;
    File(const char *path, const char *mode) { f = path ? std::fopen(path, mode) : nullptr; }
    ~File() {
        if (f) std::fclose(f);
    }
    File(const File &) = delete;
    File &operator=(const File &) = delete;
};

// sys: append the C library's reason (open/read/write failures only)
int io_error(const std::string &what, const char *path, bool sys = false) {
    std::string msg = what + ": " + (path ? path : "(null path)");
    if (sys && errno) msg += std::string(" (") + std::strerror(errno) + ")";
    sphx_set_error(msg.c_str());
    return SPHX_ERR_ARG;
}

// Split one text line at commas into doubles; returns the number of fields
// or -1 if a field is not a number.
int parse_csv_row(const char *line, double *out, int max_fields) {
    int k = 0;
    const char *p = line;
    while (*p && *p != '\n' && *p != '\r') {
        if (k >= max_fields) return max_fields + 1;
        char *end = nullptr;
        errno = 0;
        double v = std::strtod(p, &end);
        if (end == p) return -1;
        out[k++] = v;
        p = end;
        while (*p == ' ' || *p == '\t') ++p;
        if (*p == ',') ++p;
        else if (*p && *p != '\n' && *p != '\r') return -1;
    }
    return k;
}

bool read_line(FILE *f, std::string &line) {
    line.clear();
    char buf[4096];
    while (std::fgets(buf, sizeof(buf), f)) {
        line += buf;
        if (!line.empty() && line.back() == '\n') return true;
    }
    return !line.empty();
}

template <typename T>
bool write_all(FILE *f, const T *p, size_t count) {
    return std::fwrite(p, sizeof(T), count, f) == count;
}

template <typename T>
bool read_all(FILE *f, T *p, size_t count) {
    return std::fread(p, sizeof(T), count, f) == count;
}

}  // namespace

extern "C" int sphx_read_params_gpu(const char *path, sphx_params_gpu *out) {
    if (!out) { sphx_set_error("null argument"); return SPHX_ERR_ARG; }
    File F(path, "r");
    if (!F.f) return io_error("cannot open params file", path, true);
    double v[10];
    for (int k = 0; k < 10; ++k)
        if (std::fscanf(F.f, "%lf", &v[k]) != 1) return io_error("params file has fewer than 10 values", path);
    if (!(v[0] >= 0.0) || v[0] != std::floor(v[0])) return io_error("params file: N is not a count", path);
    out->N = (int64_t)v[0];
    out->c_0 = v[1];
    out->gamma = v[2];
    out->Rcld_in_pc = v[3];
    out->Rcld_in_cm = v[4];
    out->Mcld_in_g = v[5];
    out->mu = v[6];
    out->Mach = v[7];
    out->grav_const_in_cgs = v[8];
    out->G = v[9];
    return SPHX_OK;
}

extern "C" int sphx_write_params_gpu(const char *path, const sphx_params_gpu *p) {
    if (!p) { sphx_set_error("null argument"); return SPHX_ERR_ARG; }
    File F(path, "w");
    if (!F.f) return io_error("cannot create params file", path, true);
    std::fprintf(F.f, "%lld\n", (long long)p->N);
    const double v[9] = {p->c_0, p->gamma, p->Rcld_in_pc, p->Rcld_in_cm, p->Mcld_in_g,
                         p->mu, p->Mach, p->grav_const_in_cgs, p->G};
    for (double x : v) std::fprintf(F.f, "%.17g\n", x);
    return std::ferror(F.f) ? io_error("write failed", path, true) : SPHX_OK;
}

// IC or restart CSV: rows x,y,z,vx,vy,vz,m,h,eps,u[,dudt].  Reads the first
// n rows into host arrays (pos, vel: (n,3)); dudt (nullable) requires the
// 11-column restart layout.  *rows_read receives the number of data rows in
// the file (all of them are checked).
extern "C" int sphx_read_ic_csv(const char *path, int64_t n, double *pos, double *vel, double *m,
                                double *h, double *eps, double *u, double *dudt,
                                int64_t *rows_read) {
    if (n < 0 || (n > 0 && (!pos || !vel || !m || !h || !eps || !u))) {
        sphx_set_error("bad arguments");
        return SPHX_ERR_ARG;
    }
    File F(path, "r");
    if (!F.f) return io_error("cannot open IC file", path, true);
    std::string line;
    int64_t r = 0;
    int ncol = -1;
    double v[12];
    while (read_line(F.f, line)) {
        if (line.find_first_not_of(" \t\r\n") == std::string::npos) continue;
        int k = parse_csv_row(line.c_str(), v, 11);
        if (k != 10 && k != 11) {
            char msg[96];
            std::snprintf(msg, sizeof(msg), "IC row %lld: expected 10 or 11 numeric columns",
                          (long long)r + 1);
            return io_error(msg, path);
        }
        if (ncol < 0) ncol = k;
        if (k != ncol) return io_error("IC rows have different column counts", path);
        if (r < n) {
            for (int d = 0; d < 3; ++d) {
                pos[3 * r + d] = v[d];
                vel[3 * r + d] = v[3 + d];
            }
            m[r] = v[6];
            h[r] = v[7];
            eps[r] = v[8];
            u[r] = v[9];
            if (dudt) {
                if (k != 11) return io_error("dudt requested but the file has no 11th column", path);
                dudt[r] = v[10];
            }
        }
        ++r;
    }
    if (rows_read) *rows_read = r;
    if (r < n) return io_error("IC file has fewer rows than requested", path);
    return SPHX_OK;
}

// Restart CSV (no header): x,y,z,vx,vy,vz,m,h,eps,u,dudt.
extern "C" int sphx_write_restart_csv(const char *path, int64_t n, const double *pos,
                                      const double *vel, const double *m, const double *h,
                                      const double *eps, const double *u, const double *dudt) {
    if (n < 0 || (n > 0 && (!pos || !vel || !m || !h || !eps || !u))) {
        sphx_set_error("bad arguments");
        return SPHX_ERR_ARG;
    }
    File F(path, "w");
    if (!F.f) return io_error("cannot create restart file", path, true);
    for (int64_t i = 0; i < n; ++i) {
        std::fprintf(F.f, "%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g", pos[3 * i],
                     pos[3 * i + 1], pos[3 * i + 2], vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], m[i],
                     h[i], eps[i], u[i]);
        if (dudt) std::fprintf(F.f, ",%.17g", dudt[i]);
        std::fputc('\n', F.f);
    }
    return std::ferror(F.f) ? io_error("write failed", path, true) : SPHX_OK;
}

// Snapshot CSV with the reference header x,y,z,vx,vy,vz,h,rho,u.
extern "C" int sphx_write_snapshot_csv(const char *path, int64_t n, const double *pos,
                                       const double *vel, const double *h, const double *rho,
                                       const double *u) {
    if (n < 0 || (n > 0 && (!pos || !vel || !h || !rho || !u))) {
        sphx_set_error("bad arguments");
        return SPHX_ERR_ARG;
    }
    File F(path, "w");
    if (!F.f) return io_error("cannot create snapshot file", path, true);
    std::fputs("x,y,z,vx,vy,vz,h,rho,u\n", F.f);
    for (int64_t i = 0; i < n; ++i)
        std::fprintf(F.f, "%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g\n", pos[3 * i],
                     pos[3 * i + 1], pos[3 * i + 2], vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], h[i],
                     rho[i], u[i]);
    return std::ferror(F.f) ? io_error("write failed", path, true) : SPHX_OK;
}

// Binary IC (float32 payload, int32 particle types, no count in the file).
extern "C" int sphx_write_ic_bin(const char *path, int64_t n, const int32_t *typ,
                                 const double *pos, const double *vel, const double *m,
                                 const double *h, const double *eps, const double *u) {
    if (n < 0 || (n > 0 && (!pos || !vel || !m || !h || !eps || !u))) {
        sphx_set_error("bad arguments");
        return SPHX_ERR_ARG;
    }
    File F(path, "wb");
    if (!F.f) return io_error("cannot create IC file", path, true);
    std::vector<int32_t> ty(n, 0);
    if (typ) std::memcpy(ty.data(), typ, sizeof(int32_t) * n);
    bool ok = write_all(F.f, ty.data(), n);
    std::vector<float> b(n);
    auto put = [&](const double *a, int64_t stride, int off) {
        for (int64_t i = 0; i < n; ++i) b[i] = (float)a[stride * i + off];
        ok = ok && write_all(F.f, b.data(), n);
    };
    for (int d = 0; d < 3; ++d) put(pos, 3, d);
    for (int d = 0; d < 3; ++d) put(vel, 3, d);
    put(m, 1, 0);
    put(h, 1, 0);
    put(eps, 1, 0);
    put(u, 1, 0);
    return ok ? SPHX_OK : io_error("write failed", path, true);
}

extern "C" int sphx_read_ic_bin(const char *path, int64_t n, int32_t *typ, double *pos,
                                double *vel, double *m, double *h, double *eps, double *u) {
    if (n < 0 || (n > 0 && (!pos || !vel || !m || !h || !eps || !u))) {
        sphx_set_error("bad arguments");
        return SPHX_ERR_ARG;
    }
    File F(path, "rb");
    if (!F.f) return io_error("cannot open IC file", path, true);
    std::fseek(F.f, 0, SEEK_END);
    long size = std::ftell(F.f);
    std::fseek(F.f, 0, SEEK_SET);
    if (size != (long)(44 * n)) return io_error("binary IC size is not 44 bytes x N", path);
    std::vector<int32_t> ty(n);
    bool ok = read_all(F.f, ty.data(), n);
    if (typ) std::memcpy(typ, ty.data(), sizeof(int32_t) * n);
    std::vector<float> b(n);
    auto get = [&](double *a, int64_t stride, int off) {
        ok = ok && read_all(F.f, b.data(), n);
        for (int64_t i = 0; i < n; ++i) a[stride * i + off] = (double)b[i];
    };
    for (int d = 0; d < 3; ++d) get(pos, 3, d);
    for (int d = 0; d < 3; ++d) get(vel, 3, d);
    get(m, 1, 0);
    get(h, 1, 0);
    get(eps, 1, 0);
    get(u, 1, 0);
    return ok ? SPHX_OK : io_error("read failed", path);
}

// Binary snapshot: int32 N, int32 Typ[N], float32 x,y,z,vx,vy,vz,rho,h,u,mass.
extern "C" int sphx_write_snapshot_bin(const char *path, int64_t n, const int32_t *typ,
                                       const double *pos, const double *vel, const double *rho,
                                       const double *h, const double *u, const double *m) {
    if (n < 0 || n > INT32_MAX || (n > 0 && (!pos || !vel || !rho || !h || !u || !m))) {
        sphx_set_error("bad arguments");
        return SPHX_ERR_ARG;
    }
    File F(path, "wb");
    if (!F.f) return io_error("cannot create snapshot file", path, true);
    int32_t n32 = (int32_t)n;
    bool ok = write_all(F.f, &n32, 1);
    std::vector<int32_t> ty(n, 0);
    if (typ) std::memcpy(ty.data(), typ, sizeof(int32_t) * n);
    ok = ok && write_all(F.f, ty.data(), n);
    std::vector<float> b(n);
    auto put = [&](const double *a, int64_t stride, int off) {
        for (int64_t i = 0; i < n; ++i) b[i] = (float)a[stride * i + off];
        ok = ok && write_all(F.f, b.data(), n);
    };
    for (int d = 0; d < 3; ++d) put(pos, 3, d);
    for (int d = 0; d < 3; ++d) put(vel, 3, d);
    put(rho, 1, 0);
    put(h, 1, 0);
    put(u, 1, 0);
    put(m, 1, 0);
    return ok ? SPHX_OK : io_error("write failed", path, true);
}

// Reads the count first: call with n_cap = 0 (arrays may be null) to learn N.
extern "C" int sphx_read_snapshot_bin(const char *path, int64_t n_cap, int64_t *n_out,
                                      int32_t *typ, double *pos, double *vel, double *rho,
                                      double *h, double *u, double *m) {
    if (!n_out) { sphx_set_error("null argument"); return SPHX_ERR_ARG; }
    File F(path, "rb");
    if (!F.f) return io_error("cannot open snapshot file", path, true);
    int32_t n32 = 0;
    if (!read_all(F.f, &n32, 1) || n32 < 0) return io_error("bad snapshot header", path);
    int64_t n = n32;
    *n_out = n;
    std::fseek(F.f, 0, SEEK_END);
    long size = std::ftell(F.f);
    if (size != (long)(4 + 44 * n)) return io_error("binary snapshot size does not match its N", path);
    if (n_cap == 0) return SPHX_OK;
    if (n_cap < n || !pos || !vel || !rho || !h || !u || !m) {
        sphx_set_error("snapshot arrays smaller than N or null");
        return SPHX_ERR_ARG;
    }
    std::fseek(F.f, 4, SEEK_SET);
    std::vector<int32_t> ty(n);
    bool ok = read_all(F.f, ty.data(), n);
    if (typ) std::memcpy(typ, ty.data(), sizeof(int32_t) * n);
    std::vector<float> b(n);
    auto get = [&](double *a, int64_t stride, int off) {
        ok = ok && read_all(F.f, b.data(), n);
        for (int64_t i = 0; i < n; ++i) a[stride * i + off] = (double)b[i];
    };
    for (int d = 0; d < 3; ++d) get(pos, 3, d);
    for (int d = 0; d < 3; ++d) get(vel, 3, d);
    get(rho, 1, 0);
    get(h, 1, 0);
    get(u, 1, 0);
    get(m, 1, 0);
    return ok ? SPHX_OK : io_error("read failed", path);
}

// ---------------------------------------------------------------------------
// coolHeatGridZ.bin, the CHIMES table file (written by GPU_sph_Type_cooling_II/
// cooling_grids_use_EC2/step_2x.py:101-140, read by hfvGpuSphTyp_v7.cu:35-79):
// int32 n_kpc, n_nh, n_z, n_t, n_m, n_time; float32 axes kpc, nh, z, T, time;
// float32 u_evol and mu [kpc][T][nH][Z][time]; float32 metalz
// [kpc][T][nH][Z][species][time].  The file size must match the sizes.
// ---------------------------------------------------------------------------
static bool chimes_sizes(const int32_t d[6], int64_t *axes, int64_t *cells, int64_t *mcells) {
    for (int k = 0; k < 6; ++k)
        if (d[k] < (k == 4 ? 0 : 1) || d[k] > (1 << 20)) return false;
    *axes = (int64_t)d[0] + d[1] + d[2] + d[3] + d[5];
    *cells = (int64_t)d[0] * d[3] * d[1] * d[2] * d[5];
    *mcells = *cells * d[4];
    return *cells < ((int64_t)1 << 40) && *mcells < ((int64_t)1 << 40);
}

extern "C" int sphx_read_chimes(const char *path, int32_t dims[6], float *kpc, float *nh, float *z,
                                float *temp, float *time_s, float *u_evol, float *mu,
                                float *metalz) {
    if (!dims) { sphx_set_error("null argument"); return SPHX_ERR_ARG; }
    File F(path, "rb");
    if (!F.f) return io_error("cannot open CHIMES table", path, true);
    int32_t d[6];
    if (!read_all(F.f, d, 6)) return io_error("CHIMES table shorter than its header", path);
    int64_t axes, cells, mcells;
    if (!chimes_sizes(d, &axes, &cells, &mcells)) return io_error("CHIMES table header sizes out of range", path);
    std::fseek(F.f, 0, SEEK_END);
    long long size = std::ftell(F.f);
    if (size != (long long)(24 + 4 * (axes + 2 * cells + mcells)))
        return io_error("CHIMES table size does not match its header", path);
    std::memcpy(dims, d, sizeof(d));
    if (!kpc && !nh && !z && !temp && !time_s && !u_evol && !mu && !metalz) return SPHX_OK;
    if (!kpc || !nh || !z || !temp || !time_s || !u_evol || !mu || (mcells > 0 && !metalz)) {
        sphx_set_error("CHIMES arrays must all be given");
        return SPHX_ERR_ARG;
    }
    std::fseek(F.f, 24, SEEK_SET);
    bool ok = read_all(F.f, kpc, d[0]) && read_all(F.f, nh, d[1]) && read_all(F.f, z, d[2]) &&
              read_all(F.f, temp, d[3]) && read_all(F.f, time_s, d[5]) &&
              read_all(F.f, u_evol, cells) && read_all(F.f, mu, cells) &&
              (mcells == 0 || read_all(F.f, metalz, mcells));
    return ok ? SPHX_OK : io_error("read failed", path);
}

extern "C" int sphx_write_chimes(const char *path, const int32_t dims[6], const float *kpc,
                                 const float *nh, const float *z, const float *temp,
                                 const float *time_s, const float *u_evol, const float *mu,
                                 const float *metalz) {
    int64_t axes, cells, mcells;
    if (!dims || !chimes_sizes(dims, &axes, &cells, &mcells)) { sphx_set_error("bad CHIMES sizes"); return SPHX_ERR_ARG; }
    if (!kpc || !nh || !z || !temp || !time_s || !u_evol || !mu || (mcells > 0 && !metalz)) {
        sphx_set_error("null argument");
        return SPHX_ERR_ARG;
    }
    File F(path, "wb");
    if (!F.f) return io_error("cannot create CHIMES table", path, true);
    bool ok = write_all(F.f, dims, 6) && write_all(F.f, kpc, dims[0]) &&
              write_all(F.f, nh, dims[1]) && write_all(F.f, z, dims[2]) &&
              write_all(F.f, temp, dims[3]) && write_all(F.f, time_s, dims[5]) &&
              write_all(F.f, u_evol, cells) && write_all(F.f, mu, cells) &&
              (mcells == 0 || write_all(F.f, metalz, mcells));
    return ok ? SPHX_OK : io_error("write failed", path, true);
}

// ---------------------------------------------------------------------------
// Exact checkpoint of a resident run (no counterpart in the reference, whose
// restart CSV holds 6-digit floats and makes the driver recompute the forces).
// Layout (little-endian): CkptHeader, then SPHX_CKPT_ND arrays of n float64
// and SPHX_CKPT_NI arrays of n int32, every per-particle array of the engine
// in the run's internal (Morton) order, so a restored run continues
// bit-identically.
// ---------------------------------------------------------------------------
std::vector<double *> ckpt_doubles(Soa &s) {
    return {s.x,  s.y,  s.z,  s.vx,  s.vy,  s.vz,    s.u,  s.m,  s.h,  s.eps, s.ut_prev, s.ax,
            s.ay, s.az, s.gx, s.gy,  s.gz,  s.rho,   s.P,  s.c,  s.divv, s.curlv, s.du,  s.pot};
}

std::vector<int32_t *> ckpt_ints(Soa &s) { return {s.k, s.id}; }

int sphx_ckpt_write(const char *path, const CkptHeader &hdr, Engine &E, int64_t p0,
                    hipStream_t st) {
    int64_t n = hdr.n;
    File F(path, "wb");
    if (!F.f) return io_error("cannot create checkpoint", path, true);
    bool ok = write_all(F.f, &hdr, 1);
    std::vector<double> bd(n);
    for (double *a : ckpt_doubles(E.s)) {
        SPHX_HIP(hipMemcpyAsync(bd.data(), a + p0, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        SPHX_HIP(hipStreamSynchronize(st));
        ok = ok && write_all(F.f, bd.data(), n);
    }
    std::vector<int32_t> bi(n);
    for (int32_t *a : ckpt_ints(E.s)) {
        SPHX_HIP(hipMemcpyAsync(bi.data(), a + p0, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
        SPHX_HIP(hipStreamSynchronize(st));
        ok = ok && write_all(F.f, bi.data(), n);
    }
    return ok ? SPHX_OK : io_error("checkpoint write failed", path, true);
}

int sphx_ckpt_read_header(const char *path, CkptHeader *hdr) {
    File F(path, "rb");
    if (!F.f) return io_error("cannot open checkpoint", path, true);
    if (!read_all(F.f, hdr, 1)) return io_error("checkpoint header truncated", path);
    if (std::memcmp(hdr->magic, SPHX_CKPT_MAGIC, 8) != 0 || hdr->version != SPHX_CKPT_VERSION)
        return io_error("not an sphx checkpoint of this version", path);
    std::fseek(F.f, 0, SEEK_END);
    long size = std::ftell(F.f);
    long want = (long)sizeof(CkptHeader) + (long)hdr->n * (8L * SPHX_CKPT_ND + 4L * SPHX_CKPT_NI);
    if (hdr->n < 0 || size != want) return io_error("checkpoint size does not match its header", path);
    return SPHX_OK;
}

int sphx_ckpt_upload(const char *path, Engine &E, int64_t p0, int64_t n, hipStream_t st) {
    File F(path, "rb");
    if (!F.f) return io_error("cannot open checkpoint", path, true);
    std::fseek(F.f, (long)sizeof(CkptHeader), SEEK_SET);
    std::vector<double> bd(n);
    for (double *a : ckpt_doubles(E.s)) {
        if (!read_all(F.f, bd.data(), n)) return io_error("checkpoint truncated", path);
        SPHX_HIP(hipMemcpyAsync(a + p0, bd.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
        SPHX_HIP(hipStreamSynchronize(st));
    }
    std::vector<int32_t> bi(n);
    for (int32_t *a : ckpt_ints(E.s)) {
        if (!read_all(F.f, bi.data(), n)) return io_error("checkpoint truncated", path);
        SPHX_HIP(hipMemcpyAsync(a + p0, bi.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
        SPHX_HIP(hipStreamSynchronize(st));
    }
    return SPHX_OK;
}
