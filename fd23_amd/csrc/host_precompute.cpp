// Host precompute behind the fd23 C ABI: C-PML profiles, material arrays,
// moment tensors, source-time functions, model files and ISIS output.
// These keep the reference's par/model/source/receiver boundary: the
// formulas, their order of operations and their edge quirks follow the
// reference files cited at each function.
#include "fd23.h"
#include "mdl.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

void fd23_set_error_msg(const char *msg);
extern "C" void fd23_set_error_(const char *msg);

static int herr(int code, const std::string &m) {
    fd23_set_error_(m.c_str());
    return code;
}

static const double kPI = 3.14159265358979; // const.h:27

// ---------------------------------------------------------------------------
// C-PML (PSV/clean_nobox/cpml/cpml.cpp:7-83), evaluated in double.
// d is the normalised depth into the layer measured from pml_pos = npml-0.5
// on each side; sigma = sigmax d^2, kappa = 1 + d^2, alpha = pi fc (1-d);
// the first 4 and last 5 entries are zeroed afterwards (:54-64).
// ---------------------------------------------------------------------------
extern "C" int fd23_cpml_profile(int n, int npml, double startpos, double pml_dt, double pml_r, double pml_v,
                                 double pml_fc, double *b, double *c, double *k) {
    if (n <= 0 || npml < 0 || !b || !c || !k) return herr(FD23_EINVAL, "cpml_profile: bad arguments");
    const double pos = npml - 0.5;
    const double sigmax = -(2.0 + 1.0) / (2.0 * std::floor(pos)) * pml_v * std::log(pml_r);
    const double alphamax = 3.141592653589 * pml_fc;
    std::vector<double> d(n, 0.0);
    for (int i = 0; i < std::min(npml + 2, n); i++) {
        double t = (pos - (i + startpos)) / std::floor((double)npml);
        if (t > 0) d[i] = t;
    }
    for (int i = std::max(0, n - npml - 2); i < n; i++) {
        double t = ((i + startpos) - (n - 1 - pos)) / std::floor((double)npml);
        if (t > 0) d[i] = t;
    }
    for (int i = 0; i < n; i++) {
        const double sg = sigmax * d[i] * d[i];
        const double kp = 1.0 + d[i] * d[i];
        const double al = alphamax * (1.0 - d[i]);
        const double bb = std::exp(-(sg / kp + al) * pml_dt);
        b[i] = bb;
        c[i] = (sg + kp * al < 1E-15) ? 0.0 : sg * (bb - 1.0) / (sg + kp * al) / kp;
        k[i] = 1.0 / kp;
    }
    for (int i = 0; i < std::min(4, n); i++) b[i] = c[i] = k[i] = 0.0;
    for (int i = std::max(0, n - 5); i < n; i++) b[i] = c[i] = k[i] = 0.0;
    return FD23_OK;
}

// Profile order and staggering of the CPML constructors:
// P-SV PSV/clean_nobox/cpml/cpml.cpp:177-185, SH SH/cpml/cpml.cpp:172-176.
struct ProfSpec {
    bool xdir;
    double start;
};
static const ProfSpec kPsvProf[8] = {{true, 0.5}, {true, 0.0}, {true, 0.0}, {true, 0.5},
                                     {false, 0.0}, {false, 0.5}, {false, 0.0}, {false, 0.5}};
// Txx_x(0.5) Txz_x(0.0) U_x(0.0) W_x(0.5) Tzz_z(0.0) Txz_z(0.5) U_z(0.0) W_z(0.5)
static const ProfSpec kShProf[4] = {{true, 0.0}, {true, 0.5}, {false, 0.5}, {false, 0.0}};
// Tx_x(0.0) V_x(0.5) Tz_z(0.5) V_z(0.0)

extern "C" size_t fd23_cpml_pack_len(int wave, int nx, int nz) {
    if (wave == FD23_WAVE_PSV) return (size_t)3 * (4 * (size_t)nx + 4 * (size_t)nz);
    return (size_t)3 * (2 * (size_t)nx + 2 * (size_t)nz);
}

extern "C" int fd23_cpml_pack(int wave, int nx, int nz, int npml, double pml_dt, double pml_r, double pml_v,
                              double pml_fc, double *out) {
    if (!out) return herr(FD23_EINVAL, "cpml_pack: null output");
    const ProfSpec *spec = wave == FD23_WAVE_PSV ? kPsvProf : kShProf;
    const int np = wave == FD23_WAVE_PSV ? 8 : 4;
    double *p = out;
    for (int i = 0; i < np; i++) {
        const int n = spec[i].xdir ? nx : nz;
        int r = fd23_cpml_profile(n, npml, spec[i].start, pml_dt, pml_r, pml_v, pml_fc, p, p + n, p + 2 * n);
        if (r) return r;
        p += 3 * n;
    }
    return FD23_OK;
}

// ---------------------------------------------------------------------------
// P-SV material, PSV/clean_nobox/material/material.cpp:84-235 (float
// arrays with double temporaries) and its FP64 twin
// PSV/psv_clean_double/material/material.cpp.  Order: lambda/mu, water
// stabiliser in the 5-cell left/right/bottom bands, staggered averages
// (BU wraps across rows at ix=nx-1, BW = 1/rho on the last row, MUA's
// denominator repeats m2m3m4 in place of m1m2m3), Mittet free surface on
// row 0 for ix in [1, nx-2], scaling by dt/h, stability number.
// ---------------------------------------------------------------------------
template <class T>
static int psv_material(int nx, int nz, double h, double dt, const T *vp, const T *vs, const T *den, T *BU, T *BW,
                        T *MU, T *MUA, T *LAM, double *stab) {
    if (nx < 2 || nz < 2 || !vp || !vs || !den || !BU || !BW || !MU || !MUA || !LAM)
        return herr(FD23_EINVAL, "psv_material: bad arguments");
    const size_t N = (size_t)nx * nz;
    const T th = (T)h, tdt = (T)dt;
    const T fac = tdt / th;
    std::fill(BU, BU + N, T(0));
    std::fill(BW, BW + N, T(0));
    std::fill(MUA, MUA + N, T(0));
    for (size_t q = 0; q < N; q++) {
        T mu = vs[q] * vs[q] * den[q];
        LAM[q] = (T)(vp[q] * vp[q] * den[q] - 2.0 * mu);
        MU[q] = mu;
    }
    auto water = [&](size_t q) {
        if (vs[q] / vp[q] < 0.01) {
            T mu = (T)(vp[q] * vp[q] / 3.0 * den[q]);
            LAM[q] = (T)(vp[q] * vp[q] * den[q] - 2.0 * mu);
            MU[q] = mu;
        }
    };
    const int w = 5;
    for (int iz = 0; iz < nz; iz++) {
        for (int ix = 0; ix < std::min(w, nx); ix++) water((size_t)iz * nx + ix);
        for (int ix = std::max(0, nx - w); ix < nx; ix++) water((size_t)iz * nx + ix);
    }
    for (int iz = std::max(0, nz - w); iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++) water((size_t)iz * nx + ix);
    for (int iz = 1; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++) {
            const size_t q = (size_t)iz * nx + ix;
            BU[q] = (T)(2.0 / ((double)den[q] + den[q + 1 < N ? q + 1 : q]));
            BW[q] = (T)(2.0 / ((double)den[q] + den[q + nx < N ? q - nx : q]));
            if (ix + 1 < nx && iz + 1 < nz) {
                const double m1 = MU[q], m2 = MU[q + 1], m3 = MU[q - nx], m4 = MU[q + 1 - nx];
                const double dd = std::max(m2 * m3 * m4 + m1 * m3 * m4 + m1 * m2 * m4 + m2 * m3 * m4, 1E-10);
                MUA[q] = (T)(4.0 * m1 * m2 * m3 * m4 / dd);
            } else {
                MUA[q] = MU[q];
            }
        }
    for (int ix = 1; ix < nx - 1; ix++) {
        BU[ix] = (T)(4.0 / ((double)den[ix] + den[ix + 1]));
        LAM[ix] = 0;
        MU[ix] = (T)(0.5 * MU[ix]);
    }
    for (size_t q = 0; q < N; q++) {
        MU[q] *= fac;
        LAM[q] *= fac;
        BU[q] *= fac;
        BW[q] *= fac;
        MUA[q] *= fac;
    }
    double vpmax = -99999.0;
    for (size_t q = 0; q < N; q++) vpmax = std::max(vpmax, (double)vp[q]);
    const double coefsum = 1225. / 1024. + 245. / 3072. + 49. / 5120. + 5. / 7168.;
    const double s = vpmax * (double)tdt * coefsum * std::sqrt(2.0) / (double)th;
    if (stab) *stab = s;
    if (s > 1) return herr(FD23_ESTAB, "model stability number " + std::to_string(s) + " > 1");
    return FD23_OK;
}

extern "C" int fd23_psv_material(int nx, int nz, double h, double dt, const float *vp, const float *vs,
                                 const float *den, float *BU, float *BW, float *MU, float *MUA, float *LAM,
                                 double *stab) {
    return psv_material<float>(nx, nz, h, dt, vp, vs, den, BU, BW, MU, MUA, LAM, stab);
}
extern "C" int fd23_psv_material_f64(int nx, int nz, double h, double dt, const double *vp, const double *vs,
                                     const double *den, double *BU, double *BW, double *MU, double *MUA,
                                     double *LAM, double *stab) {
    return psv_material<double>(nx, nz, h, dt, vp, vs, den, BU, BW, MU, MUA, LAM, stab);
}

// SH material, SH/material/material.cpp:76-130: BV = 1/rho, harmonic
// means of mu toward +x (wrapping across rows) and -z (row 0 uses itself).
extern "C" int fd23_sh_material(int nx, int nz, double h, double dt, const float *vs, const float *den, float *BV,
                                float *MUX, float *MUZ, double *stab) {
    if (nx < 2 || nz < 2 || !vs || !den || !BV || !MUX || !MUZ) return herr(FD23_EINVAL, "sh_material: bad args");
    const size_t N = (size_t)nx * nz;
    const float fac = (float)dt / (float)h;
    auto mu_of = [&](size_t q) -> double { return vs[q] * vs[q] * den[q]; };
    for (size_t q = 0; q < N; q++) {
        BV[q] = (float)(1.0 / den[q]);
        double m1 = mu_of(q), m2 = q + 1 < N ? mu_of(q + 1) : m1;
        MUX[q] = (float)(2.0 * m1 * m2 / std::max(m1 + m2, 1E-10));
        m2 = q >= (size_t)nx ? mu_of(q - nx) : m1;
        MUZ[q] = (float)(2.0 * m1 * m2 / std::max(m1 + m2, 1E-10));
    }
    for (size_t q = 0; q < N; q++) {
        MUX[q] *= fac;
        MUZ[q] *= fac;
        BV[q] *= fac;
    }
    double vsmax = -99999.0;
    for (size_t q = 0; q < N; q++) vsmax = std::max(vsmax, (double)vs[q]);
    const double coefsum = 1225. / 1024. + 245. / 3072. + 49. / 5120. + 5. / 7168.;
    const double s = vsmax * (double)(float)dt * coefsum * std::sqrt(2.0) / (double)(float)h;
    if (stab) *stab = s;
    if (s > 1) return herr(FD23_ESTAB, "model stability number " + std::to_string(s) + " > 1");
    return FD23_OK;
}

// ---------------------------------------------------------------------------
// moment tensors, PSV/clean_nobox/point_source/cu_point_source.cu:16-40
// (kind FD23_SRC_P adds the dt/h factor, :103) and
// SH/point_source/cu_point_source.cu:16-31.  The FP32 trees round the
// angles and components to float; so does this code.
// ---------------------------------------------------------------------------
extern "C" int fd23_psv_moment(int kind, int explosion, double strike, double dip, double rake, double azimuth,
                               double h, double dt, double *M) {
    if (!M) return herr(FD23_EINVAL, "psv_moment: null output");
    float Mxx, Mxz, Mzz;
    if (explosion) {
        Mxx = 1;
        Mzz = 1;
        Mxz = 0;
    } else {
        const float delta = (float)((float)dip * kPI / 180.0);
        const float lambda = (float)((float)rake * kPI / 180.0);
        const float phi = (float)(((float)strike - (float)azimuth) * kPI / 180.0);
        Mxx = (float)(-(std::sin((double)delta) * std::cos((double)lambda) * std::sin(2.0 * phi) +
                        std::sin(2.0 * delta) * std::sin((double)lambda) * std::sin((double)phi) *
                            std::sin((double)phi)));
        Mxz = (float)(-(std::cos((double)delta) * std::cos((double)lambda) * std::cos((double)phi) +
                        std::cos(2.0 * delta) * std::sin((double)lambda) * std::sin((double)phi)));
        Mzz = (float)(std::sin(2.0 * delta) * std::sin((double)lambda));
    }
    const double hh = (float)h;
    double f = 1.0 / kPI / 1.4142135623731 / (double)(float)(hh * hh);
    if (kind == FD23_SRC_P) f = f * (float)dt / hh;
    const float factor = (float)f;
    M[0] = Mxx * factor;
    M[1] = Mxz * factor;
    M[2] = Mzz * factor;
    return FD23_OK;
}

// the FP64 tree's moment tensor (PSV/psv_clean_double/point_source/
// cu_point_source.cu:11-27, the _p factor :90): double throughout
extern "C" int fd23_psv_moment_f64(int kind, double strike, double dip, double rake, double azimuth, double h,
                                   double dt, double *M) {
    if (!M) return herr(FD23_EINVAL, "psv_moment: null output");
    const double delta = dip * kPI / 180.0, lambda = rake * kPI / 180.0, phi = (strike - azimuth) * kPI / 180.0;
    double Mxx = -(std::sin(delta) * std::cos(lambda) * std::sin(2.0 * phi) +
                   std::sin(2.0 * delta) * std::sin(lambda) * std::sin(phi) * std::sin(phi));
    double Mxz = -(std::cos(delta) * std::cos(lambda) * std::cos(phi) + std::cos(2.0 * delta) * std::sin(lambda) *
                                                                           std::sin(phi));
    double Mzz = std::sin(2.0 * delta) * std::sin(lambda);
    double factor = 1.0 / kPI / 1.4142135623731 / (h * h);
    if (kind == FD23_SRC_P) factor = factor * dt / h;
    M[0] = Mxx * factor;
    M[1] = Mxz * factor;
    M[2] = Mzz * factor;
    return FD23_OK;
}

extern "C" int fd23_sh_moment(int kind, double strike, double dip, double rake, double azimuth, double h, double dt,
                              double *M) {
    if (!M) return herr(FD23_EINVAL, "sh_moment: null output");
    const double delta = (float)dip * kPI / 180.0;
    const double lambda = (float)rake * kPI / 180.0;
    const double phi = ((float)strike - (float)azimuth) * kPI / 180.0;
    const float Mxy = (float)(std::sin(delta) * std::cos(lambda) * std::cos(2.0 * phi) +
                              0.5 * std::sin(2.0 * delta) * std::sin(lambda) * std::sin(2.0 * phi));
    const float Mzy = (float)(-(std::cos(delta) * std::cos(lambda) * std::sin(phi) -
                                std::cos(2.0 * delta) * std::sin(lambda) * std::cos(phi)));
    const double hh = (float)h;
    double f = 1.0 / kPI / 1.4142135623731 / (double)(float)(hh * hh);
    if (kind == FD23_SRC_P) f = f * (float)dt / hh;
    const float factor = (float)f;
    M[0] = Mxy * factor;
    M[1] = Mzy * factor;
    return FD23_OK;
}

// ---------------------------------------------------------------------------
// source-time functions, PSV/clean_nobox/source/waveform.cpp:8-47
// ---------------------------------------------------------------------------
extern "C" int fd23_stf_gaussian(int alpha, double *out, int cap, int *lsrc) {
    const int aint = std::abs((int)(alpha * 3.0));
    const int n = 2 * aint + 1;
    if (lsrc) *lsrc = n;
    if (!out || cap < n) return herr(FD23_EINVAL, "stf_gaussian: output too small");
    std::vector<float> s(n);
    if (alpha >= 0) {
        for (int i = -aint; i <= aint; i++)
            s[i + aint] = (float)i * (float)std::exp((double)((float)(-i) * (float)i / (float)(alpha * alpha)));
    } else {
        float sum = 0;
        for (int i = -aint; i <= aint; i++) {
            s[i + aint] = (float)std::exp((double)((float)(-i) * (float)i / (float)(alpha * alpha)));
            sum += s[i + aint];
        }
        for (int i = 0; i < n; i++) s[i] /= sum;
    }
    for (int i = 0; i < n; i++) out[i] = s[i];
    return FD23_OK;
}

// the FP64 tree's Gaussian (PSV/psv_clean_double/source/waveform.cpp:27-47):
// double arithmetic, alpha an int
extern "C" int fd23_stf_gaussian_f64(int alpha, double *out, int cap, int *lsrc) {
    const int aint = std::abs((int)(alpha * 3.0));
    const int n = 2 * aint + 1;
    if (lsrc) *lsrc = n;
    if (!out || cap < n) return herr(FD23_EINVAL, "stf_gaussian: output too small");
    if (alpha >= 0) {
        for (int i = -aint; i <= aint; i++) out[i + aint] = (double)i * std::exp((double)(-i) * (double)i / (alpha * alpha));
    } else {
        double sum = 0.0;
        for (int i = -aint; i <= aint; i++) {
            out[i + aint] = std::exp((double)(-i) * (double)i / (alpha * alpha));
            sum += out[i + aint];
        }
        for (int i = 0; i < n; i++) out[i] /= sum;
    }
    return FD23_OK;
}

// the FP64 tree's trapezoid (PSV/psv_clean_double/source/waveform.cpp:8-25)
extern "C" int fd23_stf_trapezoid_f64(double dt, double trap1, double trap2, double trap3, double *out, int cap,
                                      int *lsrc) {
    int t1 = (int)(trap1 / dt), t2 = (int)(trap2 / dt), t3 = (int)(trap3 / dt);
    t2 += t1;
    const int n = t2 + t3;
    if (lsrc) *lsrc = n;
    if (!out || cap < n) return herr(FD23_EINVAL, "stf_trapezoid: output too small");
    const double amp = (trap2 + 0.5 * (trap1 + trap3)) / dt;
    int i, j = 0;
    for (i = 0; i < t1; i++) out[j++] = i / (t1 * amp);
    for (; i < t2; i++) out[j++] = 1.0 / amp;
    for (i = t3; i > 0; i--) out[j++] = i / (t3 * amp);
    return FD23_OK;
}

extern "C" int fd23_stf_trapezoid(double dt, double trap1, double trap2, double trap3, double *out, int cap,
                                  int *lsrc) {
    const float fdt = (float)dt, f1 = (float)trap1, f2 = (float)trap2, f3 = (float)trap3;
    int t1 = (int)(f1 / fdt), t2 = (int)(f2 / fdt), t3 = (int)(f3 / fdt);
    t2 += t1;
    const int n = t2 + t3;
    if (lsrc) *lsrc = n;
    if (!out || cap < n) return herr(FD23_EINVAL, "stf_trapezoid: output too small");
    const float amp = (float)((f2 + 0.5 * (f1 + f3)) / fdt);
    int i, j = 0;
    for (i = 0; i < t1; i++) out[j++] = (float)(i / (t1 * amp));
    for (; i < t2; i++) out[j++] = (float)(1.0 / amp);
    for (i = t3; i > 0; i--) out[j++] = (float)(i / (t3 * amp));
    return FD23_OK;
}

// Ricker wavelet (1 - 2a) exp(-a), a = (pi f0 (t - t0))^2, t = i*dt; not in
// the reference, used by the synthetic workloads.
extern "C" int fd23_stf_ricker(int n, double dt, double f0, double t0, double *out) {
    if (n < 0 || !out) return herr(FD23_EINVAL, "stf_ricker: bad arguments");
    for (int i = 0; i < n; i++) {
        double a = 3.14159265358979323846 * f0 * (i * dt - t0);
        a *= a;
        out[i] = (1.0 - 2.0 * a) * std::exp(-a);
    }
    return FD23_OK;
}

// ---------------------------------------------------------------------------
// models: raster "{name}.vp/.vs/.den" (SH/model/util.cpp:12-29) or the
// ".mdl" polygon language of SH/model/model.cpp:240-369, rasterised at
// x = ix*h, z = iz*h with later polygons painted over earlier ones
// (:175-237) and the crossing-number test of :437-469.
// ---------------------------------------------------------------------------
// values and vertices as parsed (double): the FP32 trees keep them as float
// (SH/model/model.cpp:12-26), the FP64 tree as double
// (PSV/psv_clean_double/model/model.cpp:12-26); readers round as theirs do
struct Poly {
    double vp, vs, den;
    std::vector<std::pair<double, double>> v;
};

static bool ends_with(const std::string &s, const std::string &e) {
    return s.size() > e.size() && s.compare(s.size() - e.size(), e.size(), e) == 0 && s[s.size() - e.size() - 1] == '.';
}

static std::vector<std::string> split_ws(const std::string &line) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && (line[i] == ' ' || line[i] == '\t')) i++;
        if (i >= line.size()) break;
        if (line[i] == '"') {
            size_t j = line.find('"', i + 1);
            if (j == std::string::npos) j = line.size();
            out.push_back(line.substr(i + 1, j - i - 1));
            i = j + 1;
        } else {
            size_t j = i;
            while (j < line.size() && line[j] != ' ' && line[j] != '\t') j++;
            out.push_back(line.substr(i, j - i));
            i = j;
        }
    }
    return out;
}

static int load_mdl(const std::string &path, std::vector<Poly> &polys) {
    std::ifstream in(path);
    if (!in) return herr(FD23_EIO, "cannot open model file " + path);
    std::vector<std::string> lines;
    for (std::string l; std::getline(in, l);) lines.push_back(l);
    std::vector<std::pair<std::string, std::pair<double, double>>> defs;
    bool have_back = false;
    for (size_t i = 0; i < lines.size();) {
        const std::string line = lines[i++];
        if (line.empty() || line[0] == '#') continue;
        auto a = split_ws(line);
        if (a.empty()) return herr(FD23_EINVAL, "model syntax error at line " + std::to_string(i));
        std::string key = a[0];
        for (auto &ch : key) ch = (char)std::toupper((unsigned char)ch);
        if (key == "DEF") {
            if (a.size() != 4) return herr(FD23_EINVAL, "model: DEF needs 3 arguments");
            defs.push_back({a[1], {atof(a[2].c_str()), atof(a[3].c_str())}});
            continue;
        }
        if (key != "BACK" && key != "POLY") return herr(FD23_EINVAL, "bad model command: " + line);
        if (a.size() != 4) return herr(FD23_EINVAL, "model: " + key + " needs 3 arguments");
        double vp = atof(a[1].c_str()), vs = atof(a[2].c_str()), den = atof(a[3].c_str());
        if (vs < 0) vs = -vp * vs;
        if (den < 0) den = -vp * den;
        Poly p{vp, vs, den, {}};
        if (key == "BACK") {
            if (polys.empty()) polys.push_back(p);
            else polys[0] = p;
            have_back = true;
            continue;
        }
        while (i < lines.size() && !lines[i].empty() && lines[i][0] == '\t') {
            auto t = split_ws(lines[i++]);
            for (size_t j = 0; j < t.size(); j++) {
                if (std::isalpha((unsigned char)t[0][0])) {
                    bool found = false;
                    for (auto &d : defs)
                        if (d.first == t[j]) {
                            p.v.push_back(d.second);
                            found = true;
                        }
                    if (!found) return herr(FD23_EINVAL, "undefined model symbol " + t[j]);
                } else {
                    if (j + 1 >= t.size()) return herr(FD23_EINVAL, "model: odd vertex list");
                    p.v.push_back({atof(t[j].c_str()), atof(t[j + 1].c_str())});
                    j++;
                }
            }
        }
        if (polys.empty()) polys.push_back(Poly{0, 0, 0, {}});
        polys.push_back(p);
    }
    if (!have_back) return herr(FD23_EINVAL, "model has no BACK line");
    return FD23_OK;
}

template <class V> static int inside(double xp, double yp, const std::vector<std::pair<V, V>> &v) {
    const int n = (int)v.size();
    if (n == 0) return 0;
    int nleft = 0;
    double x2 = v[n - 1].first, y2 = v[n - 1].second;
    for (int j = 0; j < n; j++) {
        const double x1 = x2, y1 = y2;
        x2 = v[j].first;
        y2 = v[j].second;
        if (std::min(y1, y2) >= yp) continue;
        if (std::max(y1, y2) < yp) continue;
        if (y1 == y2) {
            if (xp > std::max(x1, x2)) continue;
            if (xp >= std::min(x1, x2)) return 1;
            nleft++;
            continue;
        }
        const double xi = x1 + (yp - y1) * (x2 - x1) / (y2 - y1);
        if (xi == xp) return 1;
        if (xi > xp) nleft++;
    }
    return nleft % 2;
}

int fd23_mdl_flatten(const char *path, int nx, int nz, double h, MdlFlat &out) {
    std::vector<Poly> polys;
    int r = load_mdl(path, polys);
    if (r) return r;
    out = MdlFlat{};
    const double hh = (float)h;
    auto clampi = [](int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); };
    for (size_t p = 0; p < polys.size(); p++) {
        const Poly &P = polys[p];
        float x1 = 1E28f, x2 = -1E28f, z1 = 1E28f, z2 = -1E28f;
        out.first.push_back((int)out.vx.size());
        out.count.push_back(p == 0 ? 0 : (int)P.v.size());
        for (auto &v : P.v) {
            if (p == 0) break;
            const float vx = (float)v.first, vz = (float)v.second;
            out.vx.push_back(vx);
            out.vz.push_back(vz);
            x1 = std::min(x1, vx);
            x2 = std::max(x2, vx);
            z1 = std::min(z1, vz);
            z2 = std::max(z2, vz);
        }
        out.ix1.push_back(clampi((int)(x1 / hh - 1), 0, nx - 1));
        out.ix2.push_back(clampi((int)(x2 / hh + 1), 0, nx - 1));
        out.iz1.push_back(clampi((int)(z1 / hh - 1), 0, nz - 1));
        out.iz2.push_back(clampi((int)(z2 / hh + 1), 0, nz - 1));
        out.vp.push_back((float)P.vp);
        out.vs.push_back((float)P.vs);
        out.den.push_back((float)P.den);
    }
    return FD23_OK;
}

extern "C" int fd23_model_read(const char *name, int nx, int nz, double h, float *vp, float *vs, float *den) {
    if (!name || nx < 1 || nz < 1 || !vp || !vs || !den) return herr(FD23_EINVAL, "model_read: bad arguments");
    const std::string nm(name);
    const size_t N = (size_t)nx * nz;
    if (ends_with(nm, "mdl")) {
        MdlFlat m;
        int r = fd23_mdl_flatten(name, nx, nz, h, m);
        if (r) return r;
        for (size_t q = 0; q < N; q++) {
            vp[q] = m.vp[0];
            vs[q] = m.vs[0];
            den[q] = m.den[0];
        }
        const double hh = (float)h;
        for (size_t p = 1; p < m.vp.size(); p++) {
            if (m.count[p] == 0) continue;
            std::vector<std::pair<float, float>> v;
            for (int j = 0; j < m.count[p]; j++) v.push_back({m.vx[m.first[p] + j], m.vz[m.first[p] + j]});
            for (int iz = m.iz1[p]; iz <= m.iz2[p]; iz++)
                for (int ix = m.ix1[p]; ix <= m.ix2[p]; ix++)
                    if (inside(hh * ix, hh * iz, v)) {
                        const size_t q = (size_t)iz * nx + ix;
                        vp[q] = m.vp[p];
                        vs[q] = m.vs[p];
                        den[q] = m.den[p];
                    }
        }
        return FD23_OK;
    }
    const char *ext[3] = {"vp", "vs", "den"};
    float *dst[3] = {vp, vs, den};
    for (int i = 0; i < 3; i++) {
        const std::string fn = nm + "." + ext[i];
        std::ifstream in(fn, std::ios::binary);
        if (!in) return herr(FD23_EIO, fn + " does not exist");
        in.read((char *)dst[i], (std::streamsize)(N * sizeof(float)));
        if ((size_t)in.gcount() != N * sizeof(float)) return herr(FD23_EIO, fn + " is shorter than nx*nz floats");
    }
    return FD23_OK;
}

// The FP64 tree's reader (PSV/psv_clean_double/model/model.cpp:175-237,
// util.cpp:12-27, material.cpp:101-150): .mdl polygons with double
// vertices, values and cell bounds; raster files are float, widened.
extern "C" int fd23_model_read_f64(const char *name, int nx, int nz, double h, double *vp, double *vs,
                                   double *den) {
    if (!name || nx < 1 || nz < 1 || !vp || !vs || !den) return herr(FD23_EINVAL, "model_read: bad arguments");
    const std::string nm(name);
    const size_t N = (size_t)nx * nz;
    if (!ends_with(nm, "mdl")) {
        std::vector<float> a(N), b(N), c(N);
        int r = fd23_model_read(name, nx, nz, h, a.data(), b.data(), c.data());
        if (r) return r;
        for (size_t q = 0; q < N; q++) {
            vp[q] = a[q];
            vs[q] = b[q];
            den[q] = c[q];
        }
        return FD23_OK;
    }
    std::vector<Poly> polys;
    int r = load_mdl(nm, polys);
    if (r) return r;
    for (size_t q = 0; q < N; q++) {
        vp[q] = polys[0].vp;
        vs[q] = polys[0].vs;
        den[q] = polys[0].den;
    }
    auto clampi = [](int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); };
    for (size_t p = 1; p < polys.size(); p++) {
        const Poly &P = polys[p];
        if (P.v.empty()) continue;
        double x1 = 1E28, x2 = -1E28, z1 = 1E28, z2 = -1E28;
        for (auto &v : P.v) {
            x1 = std::min(x1, v.first);
            x2 = std::max(x2, v.first);
            z1 = std::min(z1, v.second);
            z2 = std::max(z2, v.second);
        }
        const int ix1 = clampi((int)(x1 / h - 1), 0, nx - 1), ix2 = clampi((int)(x2 / h + 1), 0, nx - 1);
        const int iz1 = clampi((int)(z1 / h - 1), 0, nz - 1), iz2 = clampi((int)(z2 / h + 1), 0, nz - 1);
        for (int iz = iz1; iz <= iz2; iz++)
            for (int ix = ix1; ix <= ix2; ix++)
                if (inside(h * (double)ix, h * (double)iz, P.v)) {
                    const size_t q = (size_t)iz * nx + ix;
                    vp[q] = P.vp;
                    vs[q] = P.vs;
                    den[q] = P.den;
                }
    }
    return FD23_OK;
}

// ---------------------------------------------------------------------------
// ISIS writer, PSV/clean_nobox/record/isis.cpp:8-61 with the 84-byte header
// of record/isis.h:7-26.  Fields the reference leaves uninitialised
// (recorder, chan, scomp, rcomp, user[0], user[3]) are written as zero.
// ---------------------------------------------------------------------------
#pragma pack(push, 4)
struct IsisHeader {
    float xs, ys, zs, xr, yr, zr, t0;
    int nt;
    float samplerate;
    int srcflagnum, recflagnum, status;
    float gain, dtstatic, ampstatic;
    unsigned short recorder, chan, scomp, rcomp;
    int user[4];
};
#pragma pack(pop)
static_assert(sizeof(IsisHeader) == 84, "ISIS trace header is 84 bytes");

extern "C" int fd23_isis_write(const char *path, const float *raw, int nrec, int ntrecord, const int *rx,
                               const int *rz, double shiftx, double shiftz, double h, double dt, int itrecord, int xs,
                               int zs) {
    if (!path || (!raw && nrec * ntrecord > 0) || !rx || !rz || nrec < 0 || ntrecord < 0)
        return herr(FD23_EINVAL, "isis_write: bad arguments");
    std::ofstream out(path, std::ios::binary);
    if (!out) return herr(FD23_EIO, std::string("cannot create ") + path);
    const float fh = (float)h, fdt = (float)dt;
    std::vector<float> tr(ntrecord);
    for (int ir = 0; ir < nrec; ir++) {
        IsisHeader hd;
        memset(&hd, 0, sizeof hd);
        hd.xr = (float)((rx[ir] + (float)shiftx) * fh);
        hd.zr = (float)((rz[ir] + (float)shiftz) * fh);
        hd.xs = (float)xs * fh;
        hd.zs = (float)zs * fh;
        hd.nt = ntrecord;
        hd.samplerate = fdt * (float)itrecord;
        hd.gain = fdt / fh;
        hd.ampstatic = 1.0f;
        hd.srcflagnum = xs;
        hd.recflagnum = rx[ir];
        hd.user[1] = 123456789;
        hd.user[2] = 123456789;
        out.write((const char *)&hd, sizeof hd);
        for (int k = 0; k < ntrecord; k++) tr[k] = raw[(size_t)k * nrec + ir];
        out.write((const char *)tr.data(), (std::streamsize)(sizeof(float) * ntrecord));
    }
    return out ? FD23_OK : herr(FD23_EIO, std::string("write failed: ") + path);
}

// the FP64 tree's ISIS writer (PSV/psv_clean_double/record/isis.cpp:8-61,
// record/isis.h:7-26): a 136-byte header of doubles with natural
// alignment, the source position shifted by src_shift, user[1] = bw, double
// samples.  Fields and padding the reference leaves uninitialised are 0.
struct IsisHeader64 {
    double xs, ys, zs, xr, yr, zr, t0;
    int nt;
    double samplerate;
    int srcflagnum, recflagnum, status;
    double gain, dtstatic, ampstatic;
    unsigned short recorder, chan, scomp, rcomp;
    int user[4];
};
static_assert(sizeof(IsisHeader64) == 136, "FP64 ISIS trace header is 136 bytes");

extern "C" int fd23_isis_write_f64(const char *path, const double *raw, int nrec, int ntrecord, const int *rx,
                                   const int *rz, double shiftx, double shiftz, double h, double dt, int itrecord,
                                   int xs, int zs, double src_shift_x, double src_shift_z, int bw) {
    if (!path || (!raw && nrec * ntrecord > 0) || !rx || !rz || nrec < 0 || ntrecord < 0)
        return herr(FD23_EINVAL, "isis_write: bad arguments");
    std::ofstream out(path, std::ios::binary);
    if (!out) return herr(FD23_EIO, std::string("cannot create ") + path);
    std::vector<double> tr(ntrecord);
    for (int ir = 0; ir < nrec; ir++) {
        IsisHeader64 hd;
        memset(&hd, 0, sizeof hd);
        hd.xr = (rx[ir] + shiftx) * h;
        hd.zr = (rz[ir] + shiftz) * h;
        hd.xs = (double)(xs + src_shift_x) * h;
        hd.zs = (double)(zs + src_shift_z) * h;
        hd.nt = ntrecord;
        hd.samplerate = dt * (double)itrecord;
        hd.gain = dt / h;
        hd.ampstatic = 1.0;
        hd.srcflagnum = xs;
        hd.recflagnum = rx[ir];
        hd.user[1] = bw;
        hd.user[2] = 123456789;
        out.write((const char *)&hd, sizeof hd);
        for (int k = 0; k < ntrecord; k++) tr[k] = raw[(size_t)k * nrec + ir];
        out.write((const char *)tr.data(), (std::streamsize)(sizeof(double) * ntrecord));
    }
    return out ? FD23_OK : herr(FD23_EIO, std::string("write failed: ") + path);
}

// ---------------------------------------------------------------------------
// First-arrival window per block of 32 x 16 cells (SLIDE::SLIDE): the
// block's earliest and latest travel time, (min - before)/dt and
// (max + after)/dt truncated toward zero, in float for the FP32 trees
// (SH/slide/cu_slide.cu:7-61, PSV/clean_nobox/slide/cu_slide.cu) and in
// double for the FP64 tree (PSV/psv_clean_double/slide/cu_slide.cu:7-52,
// whose timefile holds doubles).  Quotients beyond the int range are
// clamped to it (the reference's int() is undefined there), NaN gives 0.
// ---------------------------------------------------------------------------
template <class T>
static int step_trunc(T v) {
    if (!(v == v)) return 0;
    if (v >= (T)2147483520.0f) return 2147483647;
    if (v <= (T)-2147483648.0) return (-2147483647 - 1);
    return (int)v;
}

template <class T>
static int punch_blocks(const T *times, int nx, int nz, T before, T after, T dt, int *first, int *last) {
    if (!times || !first || !last) return herr(FD23_EINVAL, "punch_blocks: null argument");
    if (nx <= 0 || nz <= 0 || nx % FD23_PUNCH_BX || nz % FD23_PUNCH_BZ)
        return herr(FD23_EINVAL, "a first-arrival window needs nx a multiple of 32 and nz of 16");
    if (!(dt > 0)) return herr(FD23_EINVAL, "bad dt for the first-arrival window");
    const int nxb = nx / FD23_PUNCH_BX, nzb = nz / FD23_PUNCH_BZ;
    for (int k = 0; k < nzb; k++)
        for (int i = 0; i < nxb; i++) {
            T lo = (T)1e28, hi = (T)-1e28;
            for (int kk = 0; kk < FD23_PUNCH_BZ; kk++)
                for (int ii = 0; ii < FD23_PUNCH_BX; ii++) {
                    const T t = times[(size_t)(k * FD23_PUNCH_BZ + kk) * nx + i * FD23_PUNCH_BX + ii];
                    lo = std::min(lo, t);
                    hi = std::max(hi, t);
                }
            first[(size_t)k * nxb + i] = step_trunc<T>((lo - before) / dt);
            last[(size_t)k * nxb + i] = step_trunc<T>((hi + after) / dt);
        }
    return FD23_OK;
}

extern "C" int fd23_punch_blocks(const float *times, int nx, int nz, float before, float after, float dt, int *first,
                                 int *last) {
    return punch_blocks<float>(times, nx, nz, before, after, dt, first, last);
}

extern "C" int fd23_punch_blocks_f64(const double *times, int nx, int nz, double before, double after, double dt,
                                     int *first, int *last) {
    return punch_blocks<double>(times, nx, nz, before, after, dt, first, last);
}

// ---------------------------------------------------------------------------
// receivers: every station gives the pair (x, z) and (x-1, z)
// (RECORD::init_line and init_list, PSV/clean_nobox/record/cu_record.cu:12-76)
// ---------------------------------------------------------------------------
extern "C" int fd23_line_receivers(int nsta, int ix0, int iz0, int idx, int idz, int *rx, int *rz) {
    if (nsta < 0 || (nsta > 0 && (!rx || !rz))) return herr(FD23_EINVAL, "line_receivers: bad arguments");
    for (int i = 0; i < nsta; i++) {
        rx[2 * i] = ix0 + idx * i;
        rz[2 * i] = iz0 + idz * i;
        rx[2 * i + 1] = rx[2 * i] - 1;
        rz[2 * i + 1] = rz[2 * i];
    }
    return FD23_OK;
}

// station file of "z x" integer pairs, at most 100000 stations
extern "C" int fd23_station_read(const char *path, int *rx, int *rz, int cap, int *nrec) {
    if (!path || !nrec) return herr(FD23_EINVAL, "station_read: bad arguments");
    std::ifstream f(path);
    if (!f.is_open()) return herr(FD23_EIO, std::string(path) + " does not exist");
    std::vector<int> v;
    int a;
    while (f >> a) v.push_back(a);
    if (!f.eof()) return herr(FD23_EIO, std::string("non-integer entry in ") + path);
    if (v.empty() || v.size() % 2) return herr(FD23_EIO, std::string("station file needs 'z x' pairs: ") + path);
    const int nsta = (int)(v.size() / 2);
    if (nsta > 100000) return herr(FD23_EINVAL, "too many stations (limit 100000)");
    *nrec = 2 * nsta;
    if (!rx || !rz) return FD23_OK; // size query
    if (cap < 2 * nsta) return herr(FD23_EINVAL, "station_read: capacity too small");
    for (int i = 0; i < nsta; i++) {
        rz[2 * i] = v[2 * i];
        rx[2 * i] = v[2 * i + 1];
        rx[2 * i + 1] = rx[2 * i] - 1;
        rz[2 * i + 1] = rz[2 * i];
    }
    return FD23_OK;
}
