// Command-line driver with the reference programs' interface, on top of the
// C ABI only.  One executable; the name it is started under picks the
// reference build variant (PSV/clean_nobox/Makefile, SH/Makefile,
// PSV/psv_clean_double/Makefile):
//   nbpsv2d, nbpsv2d_topabsorb, nbpsv2d_exp, nbpsv2d_exp_topabsorb, sh2d,
//   sh2d_topabsorb, psv2d_db, psv2d_db_topabsorb (FP64: double parameters,
//   model, wavelet and moment, {output}_U.db.isis); under any other name
//   the par keys wave=psv|sh, topabsorb=0|1 and explosion=0|1 choose.
//
// P-SV follows PSV/clean_nobox/main.cpp:24-330 and param/param.cpp:94-193;
// SH follows SH/main.cpp:20-200 and SH/param/param.cpp:12-84:
//   model (.mdl polygons or raster files) -> material -> C-PML -> source
//   wavelet (Gaussian or trapezoid; running sum for srctype=p / stype=13)
//   -> receivers (a line or station files, pairs x and x-1) -> optional
//   restart from state%06d -> time loop (graph-replayed chunks that end at
//   each snapshot and progress-print step) -> optional state%06d ->
//   raw time-major seismograms {output}_U/_W or _V -> ISIS files.
// Keys the one-GPU build reads but does not need: gpuid, ngpu, usetable.
// psv2d_db usebox=1 (srcden, srcsvel, srcpvel, sourcefile, bw) drives the
// reference's box from its sourcefile frames (fd23_box_configure_frames).
// snapbox (PSV/clean_nobox_new) writes windowed snapshots to ./OUTPUT/.
// Additional key: storage=f32|f16|bf16|f64 (default f32).  usepunch=1
// reads before, after and timefile (an nz*nx travel-time grid, float32, or
// double for psv2d_db as in the FP64 tree) and
// sets the first-arrival window (SH/slide/cu_slide.cu:7-61).
#include "fd23.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iomanip>
#include <iterator>
#include <set>
#include <string>
#include <vector>

#include <sys/stat.h>

namespace {

const int kGetparError = 100;

struct Par {
    fd23_par *p = nullptr;
    ~Par() {
        if (p) fd23_par_close(p);
    }
    bool has_i(const char *k, int &v) { return fd23_par_int(p, k, &v) == 1; }
    bool has_f(const char *k, float &v) { return fd23_par_float(p, k, &v) == 1; }
    bool has_d(const char *k, double &v) { return fd23_par_double(p, k, &v) == 1; }
    bool has_s(const char *k, std::string &v) {
        char buf[1024];
        if (fd23_par_string(p, k, buf, sizeof buf) != 1) return false;
        v = buf;
        return true;
    }
    // mstpar: a missing key ends the program, as getpar's mstpar does
    int mi(const char *k) {
        int v;
        if (!has_i(k, v)) missing(k);
        return v;
    }
    float mf(const char *k) {
        float v;
        if (!has_f(k, v)) missing(k);
        return v;
    }
    double md(const char *k) {
        double v;
        if (!has_d(k, v)) missing(k);
        return v;
    }
    std::string ms(const char *k) {
        std::string v;
        if (!has_s(k, v)) missing(k);
        return v;
    }
    // getpar reports a missing key through gp_getpar_err, which exits with
    // GETPAR_ERROR = 100 (PSV/clean_nobox/getpar/getpar.c:82,442,724-734)
    [[noreturn]] static void missing(const char *k) {
        fprintf(stderr, "mstpar: must specify value for '%s'\n", k);
        exit(kGetparError);
    }
};

[[noreturn]] void die(const char *what, int code) {
    const char *e = fd23_last_error();
    fprintf(stderr, "%s failed (%d): %s\n", what, code, e ? e : "");
    exit(-1);
}
#define CHECK(call)                                                                                                    \
    do {                                                                                                               \
        int rc_ = (call);                                                                                              \
        if (rc_ != FD23_OK) die(#call, rc_);                                                                           \
    } while (0)

void write_file(const std::string &name, const void *data, size_t bytes) {
    std::ofstream f(name, std::ios::binary);
    f.write((const char *)data, (std::streamsize)bytes);
    if (!f) {
        fprintf(stderr, "cannot write %s\n", name.c_str());
        exit(-1);
    }
}

std::string basename_of(const char *p) {
    std::string s(p ? p : "");
    const size_t k = s.find_last_of('/');
    return k == std::string::npos ? s : s.substr(k + 1);
}

// SH needsnap: extra snapshot steps listed in a file (SH/record/needsnap.cpp)
std::set<int> snap_list(const std::string &file) {
    std::set<int> s;
    if (file.empty()) return s;
    std::ifstream f(file);
    if (!f.is_open()) {
        fprintf(stderr, "%s Not exist!", file.c_str());
        exit(-1);
    }
    int v;
    while (f >> v) s.insert(v);
    return s;
}

} // namespace

int main(int argc, char **argv) {
    const std::string prog = basename_of(argv[0]);
    Par par;
    {
        // a par file that cannot be opened is a getpar error too (getpar.c:646)
        const int rc = fd23_par_open(argc, (const char *const *)argv, &par.p);
        if (rc != FD23_OK) {
            const char *e = fd23_last_error();
            fprintf(stderr, "getpar: %s\n", e ? e : "cannot parse the command line");
            exit(kGetparError);
        }
    }

    bool psv = true, top = false, explosion = false;
    // the FP64 tree (PSV/psv_clean_double/Makefile: psv2d_db[_topabsorb]):
    // its par keys (param.cpp:64-154), double parameters, double host steps
    // and the .db.isis writer
    const bool dbl = prog.rfind("psv2d_db", 0) == 0;
    if (dbl) {
        top = prog.find("topabsorb") != std::string::npos;
    } else if (prog.rfind("nbpsv2d", 0) == 0 || prog.rfind("sh2d", 0) == 0) {
        psv = prog.rfind("nbpsv2d", 0) == 0;
        top = prog.find("topabsorb") != std::string::npos;
        explosion = prog.find("_exp") != std::string::npos;
    } else {
        std::string w;
        if (par.has_s("wave", w)) psv = !(w == "sh" || w == "SH");
        int v;
        if (par.has_i("topabsorb", v)) top = v != 0;
        if (par.has_i("explosion", v)) explosion = v != 0;
    }
    int storage = FD23_STORE_F32;
    {
        std::string st;
        if (par.has_s("storage", st)) {
            if (st == "f16") storage = FD23_STORE_F16;
            else if (st == "bf16") storage = FD23_STORE_BF16;
            else if (st == "f64") storage = FD23_STORE_F64;
            else if (st != "f32") Par::missing("storage (f32|f16|bf16|f64)");
        }
        if (dbl) storage = FD23_STORE_F64;
    }

    // ---- parameters (P-SV param.cpp:94-193, SH param.cpp:12-84) --------
    int restart = 0, nt0 = 0, savestate = 0, itprint = 100, itrecord = 1, ntsnap;
    int usepunch = 0;
    int v;
    // real parameters: float in the FP32 trees, double in the FP64 one
    auto real = [&](const char *k) { return dbl ? par.md(k) : (double)par.mf(k); };
    int usebox = 0;
    if (dbl) {
        usebox = par.mi("usebox");
        usepunch = par.mi("usepunch");
        par.mi("ngpu");
    } else if (psv) {
        par.has_i("restart", restart);
        if (restart) nt0 = par.mi("nt0");
        par.has_i("savestate", savestate);
    } else {
        usepunch = par.mi("usepunch");
        par.mi("ngpu");
    }
    const int nx = par.mi("nx"), nz = par.mi("nz"), nt = par.mi("nt");
    const double h = real("h"), dt = real("dt");
    // the FP64 tree's box keys (PSV/psv_clean_double/param/param.cpp:79-85)
    double srcden = 0, srcsvel = 0, srcpvel = 0;
    std::string sourcefile;
    int bw = 0;
    if (usebox == 1) {
        srcden = real("srcden");
        srcsvel = real("srcsvel");
        srcpvel = real("srcpvel");
        sourcefile = par.ms("sourcefile");
        bw = par.mi("bw");
    }
    char srctype = '1';
    int stype = 3;
    double src_shift_x = 0.25, src_shift_z = -0.25; // FP64 tree, param.cpp:87-91
    std::string s;
    if (dbl) {
        par.has_i("stype", stype);
        srctype = stype == 13 ? 'p' : '1';
        par.has_d("src_shift_x", src_shift_x);
        par.has_d("src_shift_z", src_shift_z);
    } else if (psv) {
        if (par.has_s("srctype", s) && !s.empty()) srctype = s[0];
    } else {
        par.has_i("stype", stype);
    }
    const std::string srctime = par.ms(psv && !dbl ? "srctime" : "sourcetime");
    double trap1 = 0, trap2 = 0, trap3 = 0, alpha = 0;
    if (srctime[0] == 'T' || srctime[0] == 't') {
        trap1 = real("trap1");
        trap2 = real("trap2");
        trap3 = real("trap3");
    }
    if (srctime[0] == 'G' || srctime[0] == 'g') alpha = real("alpha");
    const double strike = real("strike"), dip = real("dip"), rake = real("rake"), azimuth = real("azimuth");
    const int xs = par.mi("xs"), zs = par.mi("zs");
    if (usebox == 1 && (zs < bw / 2 + 4 || nz - zs < bw / 2 + 4)) { // PARAM::make_plan, param.cpp:179-182
        fprintf(stderr, "zs %d is too near to the boundary\n", zs);
        return -1;
    }
    par.has_i("itprint", itprint);
    const int npml = par.mi("npml");
    const double pml_dt = real("pml_dt"), pml_v = real("pml_v"), pml_r = real("pml_r"), pml_fc = real("pml_fc");
    if (psv && !dbl)
        par.has_i("itrecord", itrecord);
    else
        itrecord = par.mi("itrecord");
    const std::string output = par.ms("output");
    // receivers per component
    std::vector<std::vector<int>> rx(psv ? 2 : 1), rz(psv ? 2 : 1);
    std::string sta[2];
    bool use_sta = false;
    if (psv) {
        par.has_s("U_sta_file", sta[0]);
        par.has_s("W_sta_file", sta[1]);
        use_sta = !(sta[0].empty() && sta[1].empty());
    } else {
        par.has_s("V_sta_file", sta[0]);
        use_sta = !sta[0].empty();
    }
    for (size_t c = 0; c < rx.size(); c++) {
        if (use_sta) {
            int n = 0;
            CHECK(fd23_station_read(sta[c].c_str(), nullptr, nullptr, 0, &n));
            rx[c].resize(n);
            rz[c].resize(n);
            CHECK(fd23_station_read(sta[c].c_str(), rx[c].data(), rz[c].data(), n, &n));
        }
    }
    if (!use_sta) {
        const int nrec = par.mi("nrec"), ixrec0 = par.mi("ixrec0");
        int iz0[2];
        if (psv) {
            iz0[0] = par.mi("izrec0_u");
            iz0[1] = par.mi("izrec0_w");
        } else {
            iz0[0] = par.mi("izrec0_v");
        }
        const int idxrec = par.mi("idxrec"), idzrec = par.mi("idzrec");
        for (size_t c = 0; c < rx.size(); c++) {
            rx[c].resize(2 * nrec);
            rz[c].resize(2 * nrec);
            CHECK(fd23_line_receivers(nrec, ixrec0, iz0[c], idxrec, idzrec, rx[c].data(), rz[c].data()));
        }
    }
    const std::string model = par.ms("model");
    if (psv && !dbl) par.has_i("usetable", v);
    ntsnap = par.mi("ntsnap");
    std::string snapfile;
    if (!psv) par.has_s("ntsnapfile", snapfile);
    if (psv && !dbl) par.has_i("usepunch", usepunch);
    double before = 0, after = 0; // float in the FP32 trees, double in the FP64 one
    std::string timefile;
    if (usepunch == 1) { // P-SV param.cpp:187-191, SH param.cpp:80-84, FP64 param.cpp:149-153
        before = real("before");
        after = real("after");
        timefile = par.ms("timefile");
    }
    const std::set<int> extra_snaps = snap_list(snapfile);
    // snapbox=x1,x2,z1,z2,it1,it2 (PSV/clean_nobox_new/param/param.cpp:184-198,
    // record/record.cpp:27-60): snapshots of that window only, for steps in
    // [it1, it2], written to ./OUTPUT/
    int snapbox[6] = {0, nx - 1, 0, nz - 1, 0, nt};
    bool use_snapbox = false;
    if (psv && !dbl && par.has_s("snapbox", s)) {
        if (sscanf(s.c_str(), "%d,%d,%d,%d,%d,%d", &snapbox[0], &snapbox[1], &snapbox[2], &snapbox[3], &snapbox[4],
                   &snapbox[5]) != 6) {
            fprintf(stderr, "snapbox should be six integer");
            return -1;
        }
        if (snapbox[0] < 0 || snapbox[1] >= nx || snapbox[0] > snapbox[1] || snapbox[2] < 0 || snapbox[3] >= nz ||
            snapbox[2] > snapbox[3]) {
            fprintf(stderr, "snapbox %s is outside the %dx%d grid\n", s.c_str(), nx, nz);
            return -1;
        }
        use_snapbox = true;
        mkdir("OUTPUT", 0777);
    }

    // ---- model, material, C-PML -----------------------------------------
    const size_t N = (size_t)nx * nz;
    // the travel-time grid, nz rows of nx values: float32 in the FP32 trees
    // (SH/slide/cu_slide.cu:24-39), double in the FP64 tree
    // (PSV/psv_clean_double/slide/cu_slide.cu:24-38)
    std::vector<float> times;
    std::vector<double> times_d;
    if (usepunch == 1) {
        const size_t es = dbl ? sizeof(double) : sizeof(float);
        if (dbl)
            times_d.resize(N);
        else
            times.resize(N);
        std::ifstream f(timefile, std::ios::binary);
        f.read(dbl ? (char *)times_d.data() : (char *)times.data(), (std::streamsize)(N * es));
        if (!f) {
            fprintf(stderr, "cannot read %zu travel times from %s\n", N, timefile.c_str());
            return -1;
        }
    }
    std::vector<float> vp(N), vs(N), den(N);
    std::vector<double> vpd, vsd, dend; // the FP64 tree's model, read in double
    if (dbl) {
        vpd.resize(N);
        vsd.resize(N);
        dend.resize(N);
        CHECK(fd23_model_read_f64(model.c_str(), nx, nz, h, vpd.data(), vsd.data(), dend.data()));
    } else {
        CHECK(fd23_model_read(model.c_str(), nx, nz, h, vp.data(), vs.data(), den.data()));
    }
    // SH output_material=1 (SH/param/param.cpp:12-13, SH/material/
    // material.cpp:50-67): a .mdl model's vs and den as {model}_VS, _DEN,
    // two int sizes then nx*nz floats
    if (!psv && par.has_i("output_material", v) && v == 1 && model.size() > 4 &&
        model.compare(model.size() - 4, 4, ".mdl") == 0) {
        const std::pair<const char *, const std::vector<float> *> outs[2] = {{"_VS", &vs}, {"_DEN", &den}};
        for (auto &o : outs) {
            std::vector<char> buf(2 * sizeof(int) + N * sizeof(float));
            memcpy(buf.data(), &nx, sizeof(int));
            memcpy(buf.data() + sizeof(int), &nz, sizeof(int));
            memcpy(buf.data() + 2 * sizeof(int), o.second->data(), N * sizeof(float));
            write_file(model + o.first, buf.data(), buf.size());
        }
    }
    const bool f64 = storage == FD23_STORE_F64;
    const int nmat = psv ? 5 : 3;
    std::vector<std::vector<float>> matf(nmat, std::vector<float>(N));
    std::vector<std::vector<double>> matd;
    double stab = 0;
    int mr;
    if (psv && f64) {
        matd.assign(nmat, std::vector<double>(N));
        if (!dbl) { // storage=f64 on an FP32 tree's inputs: the float model widened
            vpd.assign(vp.begin(), vp.end());
            vsd.assign(vs.begin(), vs.end());
            dend.assign(den.begin(), den.end());
        }
        mr = fd23_psv_material_f64(nx, nz, h, dt, vpd.data(), vsd.data(), dend.data(), matd[0].data(),
                                   matd[1].data(), matd[2].data(), matd[3].data(), matd[4].data(), &stab);
    } else if (psv) {
        mr = fd23_psv_material(nx, nz, h, dt, vp.data(), vs.data(), den.data(), matf[0].data(), matf[1].data(),
                               matf[2].data(), matf[3].data(), matf[4].data(), &stab);
    } else {
        mr = fd23_sh_material(nx, nz, h, dt, vs.data(), den.data(), matf[0].data(), matf[1].data(), matf[2].data(),
                              &stab);
    }
    if (mr == FD23_ESTAB) {
        fprintf(stderr, "unstable: stability number %g > 1\n", stab);
        return -1;
    }
    if (mr) die("material", mr);
    if (!psv && f64) { // SH material is computed in float; widen it
        matd.assign(nmat, std::vector<double>(N));
        for (int m = 0; m < nmat; m++)
            for (size_t q = 0; q < N; q++) matd[m][q] = matf[m][q];
    }
    std::vector<const void *> matp(nmat);
    for (int m = 0; m < nmat; m++) matp[m] = f64 ? (const void *)matd[m].data() : (const void *)matf[m].data();
    const int wave = psv ? FD23_WAVE_PSV : FD23_WAVE_SH;
    std::vector<double> cpml(fd23_cpml_pack_len(wave, nx, nz));
    CHECK(fd23_cpml_pack(wave, nx, nz, npml, pml_dt, pml_r, pml_v, pml_fc, cpml.data()));

    // ---- source ----------------------------------------------------------
    std::vector<double> stf(100000);
    int lsrc = 0;
    if ((srctime[0] == 'T' || srctime[0] == 't') && dbl)
        CHECK(fd23_stf_trapezoid_f64(dt, trap1, trap2, trap3, stf.data(), (int)stf.size(), &lsrc));
    else if (srctime[0] == 'T' || srctime[0] == 't')
        CHECK(fd23_stf_trapezoid(dt, trap1, trap2, trap3, stf.data(), (int)stf.size(), &lsrc));
    else if ((srctime[0] == 'G' || srctime[0] == 'g') && dbl)
        CHECK(fd23_stf_gaussian_f64((int)alpha, stf.data(), (int)stf.size(), &lsrc));
    else if (srctime[0] == 'G' || srctime[0] == 'g')
        CHECK(fd23_stf_gaussian((int)alpha, stf.data(), (int)stf.size(), &lsrc));
    else {
        fprintf(stderr, "srctime must be Gaussian or Trapezoid\n");
        return -1;
    }
    stf.resize(lsrc);
    int kind;
    if (psv) {
        if (srctype == '1')
            kind = FD23_SRC_LINE;
        else if (srctype == 'p' || srctype == 'P')
            kind = FD23_SRC_P;
        else {
            fprintf(stderr, "Wrong srctype! Specify 1 or p.");
            return -1;
        }
    } else {
        kind = stype == 13 ? FD23_SRC_P : FD23_SRC_LINE; // other stypes fall back to 3 (SH/main.cpp:104-112)
    }
    if (kind == FD23_SRC_P && dbl) { // running sum in double (PSV/psv_clean_double/main.cpp:105-110)
        double sum = 0;
        for (auto &x : stf) {
            sum += x;
            x = sum;
        }
    } else if (kind == FD23_SRC_P) { // running sum in float (main.cpp:94-99)
        float sum = 0;
        for (auto &x : stf) {
            sum += (float)x;
            x = sum;
        }
    }
    double M[3];
    if (dbl)
        CHECK(fd23_psv_moment_f64(kind, strike, dip, rake, azimuth, h, dt, M));
    else if (psv)
        CHECK(fd23_psv_moment(kind, explosion ? 1 : 0, strike, dip, rake, azimuth, h, dt, M));
    else
        CHECK(fd23_sh_moment(kind, strike, dip, rake, azimuth, h, dt, M));

    // box source: nt+1 displacement frames of U then W on (bw+8)^2 cells,
    // read in full before any device work (the reference reads one per step
    // and would reuse a stale frame past the end of the file; a short file
    // is an error here)
    std::vector<double> frames;
    if (usebox == 1) {
        const size_t fsz = (size_t)2 * (bw + 8) * (bw + 8);
        frames.resize(fsz * (size_t)(nt + 1));
        std::ifstream f(sourcefile, std::ios::binary);
        f.read((char *)frames.data(), (std::streamsize)(frames.size() * sizeof(double)));
        if (!f) {
            fprintf(stderr, "cannot read %d box source frames of %zu doubles from %s\n", nt + 1, fsz,
                    sourcefile.c_str());
            return -1;
        }
    }

    // ---- context ----------------------------------------------------------
    const int first_frame = (nt0 + itrecord - 1) / itrecord;
    const int end_frame = (nt0 + nt + itrecord - 1) / itrecord;
    int max_rec = 0;
    for (auto &r : rx) max_rec = std::max<int>(max_rec, (int)r.size());
    fd23_desc d{};
    d.wave = wave;
    d.storage = storage;
    d.nx = nx;
    d.nz = nz;
    d.npml = npml;
    d.topabsorb = top ? 1 : 0;
    d.nshots = 1;
    d.max_frames = end_frame;
    d.max_rec = max_rec;
    fd23_ctx *ctx = nullptr;
    CHECK(fd23_ctx_create(&d, &ctx));
    CHECK(fd23_set_material(ctx, matp.data()));
    CHECK(fd23_set_cpml(ctx, cpml.data()));
    if (usebox == 1) {
        // the box replaces the point source (PSV/psv_clean_double/main.cpp:153-200)
        CHECK(fd23_box_configure_frames(ctx, xs, zs, bw, frames.data(), nt + 1, srcpvel, srcsvel, srcden, h, dt));
        const int nf = fd23_box_forced_cells(ctx);
        if (nf > 0) fprintf(stderr, "No homogenous region: %d box cells set to the source medium\n", nf);
    } else {
        CHECK(fd23_set_source(ctx, 0, kind, xs, zs, stf.data(), lsrc, M));
    }
    for (size_t c = 0; c < rx.size(); c++)
        CHECK(fd23_set_receivers(ctx, 0, (int)c, rx[c].data(), rz[c].data(), (int)rx[c].size()));
    if (usepunch == 1 && dbl) CHECK(fd23_set_punch_f64(ctx, times_d.data(), before, after, dt));
    if (usepunch == 1 && !dbl) CHECK(fd23_set_punch(ctx, times.data(), (float)before, (float)after, (float)dt));

    const size_t sbytes = fd23_state_bytes(ctx);
    if (restart) {
        char name[64];
        snprintf(name, sizeof name, "state%06d", nt0 - 1);
        std::ifstream f(name, std::ios::binary);
        std::vector<char> buf(sbytes);
        f.read(buf.data(), (std::streamsize)sbytes);
        if (!f) {
            fprintf(stderr, "cannot read %s\n", name);
            return -1;
        }
        CHECK(fd23_set_state(ctx, 0, buf.data(), sbytes));
    }

    // ---- time loop -------------------------------------------------------
    const int csize = f64 ? 8 : 4;
    std::vector<char> snap(N * csize);
    const char *snames[2] = {psv ? "U" : "V", "W"};
    const int sfields[2] = {psv ? FD23_F_U : FD23_F_V, FD23_F_W};
    const auto t0 = std::chrono::steady_clock::now();
    int it = nt0;
    while (it < nt0 + nt) {
        // run through the next step that snapshots or prints
        int e = it;
        while (e < nt0 + nt - 1 && !((e % ntsnap == 0 || extra_snaps.count(e)) && e != 0) && e % itprint != 0) e++;
        CHECK(fd23_run(ctx, it, e - it + 1, itrecord));
        if ((e % ntsnap == 0 || extra_snaps.count(e)) && e != 0) {
            CHECK(fd23_sync(ctx));
            std::cout << e << std::endl;
            for (int k = 0; k < (psv ? 2 : 1) && (!use_snapbox || (e >= snapbox[4] && e <= snapbox[5])); k++) {
                CHECK(fd23_get_field(ctx, 0, sfields[k], snap.data()));
                char name[48];
                if (!use_snapbox) {
                    snprintf(name, sizeof name, "%s%05d", snames[k], e);
                    write_file(name, snap.data(), N * csize);
                    continue;
                }
                // rows z1..z2 of columns x1..x2
                const size_t w = (size_t)(snapbox[1] - snapbox[0] + 1) * csize;
                std::vector<char> win(w * (snapbox[3] - snapbox[2] + 1));
                for (int z = snapbox[2]; z <= snapbox[3]; z++)
                    memcpy(win.data() + (size_t)(z - snapbox[2]) * w,
                           snap.data() + ((size_t)z * nx + snapbox[0]) * csize, w);
                snprintf(name, sizeof name, "./OUTPUT/%s%05d", snames[k], e);
                write_file(name, win.data(), win.size());
            }
        }
        if (e % itprint == 0) {
            CHECK(fd23_sync(ctx));
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::cout << std::setw(10) << e << " TOTAL " << std::setw(10) << el << " sec" << std::setw(10)
                      << el / (e - nt0 + 1) * 1000.0 << " msec per step\n";
            std::cout.flush();
        }
        it = e + 1;
    }
    CHECK(fd23_sync(ctx));

    if (savestate) {
        std::vector<char> buf(sbytes);
        CHECK(fd23_get_state(ctx, 0, buf.data(), sbytes));
        char name[64];
        snprintf(name, sizeof name, "state%06d", nt + nt0 - 1);
        write_file(name, buf.data(), sbytes);
    }

    // ---- seismograms: raw time-major files, then ISIS ---------------------
    const int nframes = end_frame - first_frame;
    const char *comp[2] = {psv ? "_U" : "_V", "_W"};
    const double shx[2] = {psv ? 0.5 : 0.0, 0.0}, shz[2] = {0.0, psv ? -0.5 : 0.0};
    for (size_t c = 0; c < rx.size(); c++) {
        const int n = (int)rx[c].size();
        std::vector<char> all((size_t)end_frame * n * csize);
        if (n) CHECK(fd23_get_seismograms(ctx, 0, (int)c, (float *)all.data(), end_frame));
        const std::string raw = output + comp[c];
        write_file(raw, all.data() + (size_t)first_frame * n * csize, (size_t)nframes * n * csize);
        // toisis reads floor(nt/itrecord) frames of the raw file (isis.cpp:20)
        const int ntrec = std::min(nt / itrecord, nframes);
        if (dbl) { // double samples behind 136-byte headers, {raw}.db.isis
            CHECK(fd23_isis_write_f64((raw + ".db.isis").c_str(),
                                      (const double *)all.data() + (size_t)first_frame * n, n, ntrec, rx[c].data(),
                                      rz[c].data(), shx[c], shz[c], h, dt, itrecord, xs, zs, src_shift_x,
                                      src_shift_z, 0));
            continue;
        }
        std::vector<float> tr((size_t)ntrec * n);
        for (size_t q = 0; q < tr.size(); q++)
            tr[q] = f64 ? (float)((const double *)all.data())[(size_t)first_frame * n + q]
                        : ((const float *)all.data())[(size_t)first_frame * n + q];
        CHECK(fd23_isis_write((raw + ".isis").c_str(), tr.data(), n, ntrec, rx[c].data(), rz[c].data(), shx[c],
                              shz[c], h, dt, itrecord, xs, zs));
    }
    fd23_ctx_destroy(ctx);
    std::cout << "finish\n";
    return 0;
}
