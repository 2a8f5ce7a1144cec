// Runs one pass of the reference's SH CPU loop (SH/fd/fd.cpp step_stress /
// step_velocity) on fields, psi and material read from files, with the
// reference's own C-PML profiles (CPML::CPML), and dumps the results.
// The loop differs from the GPU kernels only near the domain edges (order
// reduction without iz, no early return, row-major wrap), so the caller
// compares interior cells.  Test infrastructure.
#include "harness_common.h"
#include "param/param.h"
#include "material/material.h"
#include "field/field.h"
#include "cpml/cpml.h"
#include "fd/fd.h"
#include <cstring>
#include <map>
#include <string>
#include <vector>

static std::vector<float> load(const char *name, long n) {
    std::vector<float> v(n);
    std::string path = g_outdir + "/" + name;
    FILE *f = fopen(path.c_str(), "rb");
    if (!f || fread(v.data(), sizeof(float), (size_t)n, f) != (size_t)n) {
        fprintf(stderr, "cannot read %s\n", path.c_str());
        exit(3);
    }
    fclose(f);
    return v;
}

int main(int ac, char **av) {
    init_outdir();
    std::map<std::string, std::string> kv;
    for (int i = 1; i < ac; i++) {
        const char *e = strchr(av[i], '=');
        if (e) kv[std::string(av[i], e - av[i])] = e + 1;
    }
    auto I = [&](const char *k) { return atoi(kv.at(k).c_str()); };
    auto Fl = [&](const char *k) { return (float)atof(kv.at(k).c_str()); };
    PARAM param;
    memset(&param, 0, sizeof param);
    param.nx = I("nx");
    param.nz = I("nz");
    param.npml = I("npml");
    param.pml_dt = Fl("pml_dt");
    param.pml_r = Fl("pml_r");
    param.pml_v = Fl("pml_v");
    param.pml_fc = Fl("pml_fc");
    const int nx = param.nx, nz = param.nz, npml = param.npml;
    const long N = (long)nx * nz, NX = 2L * npml * nz, NZ = 2L * npml * nx;
    FIELD fld;
    fld.init_for_full(param);
    std::vector<float> V = load("V.in", N), Tx = load("Tx.in", N), Tz = load("Tz.in", N);
    memcpy(fld.V, V.data(), N * 4);
    memcpy(fld.Tx, Tx.data(), N * 4);
    memcpy(fld.Tz, Tz.data(), N * 4);
    std::vector<float> BV = load("BV.in", N), MUX = load("MUX.in", N), MUZ = load("MUZ.in", N);
    MATERIAL mat;
    mat.nx = nx;
    mat.nz = nz;
    mat.BV = BV.data();
    mat.MUX = MUX.data();
    mat.MUZ = MUZ.data();
    CPML cp(param);
    std::vector<float> pvx = load("psi_V_x.in", NX), pvz = load("psi_V_z.in", NZ);
    std::vector<float> ptx = load("psi_Tx_x.in", NX), ptz = load("psi_Tz_z.in", NZ);
    memcpy(cp.psi.V_x, pvx.data(), NX * 4);
    memcpy(cp.psi.V_z, pvz.data(), NZ * 4);
    memcpy(cp.psi.Tx_x, ptx.data(), NX * 4);
    memcpy(cp.psi.Tz_z, ptz.data(), NZ * 4);
    const std::string mode = kv.at("mode");
    if (mode == "stress") {
        step_stress(fld, mat, cp);
        dump("Tx.out", fld.Tx, N);
        dump("Tz.out", fld.Tz, N);
        dump("psi_V_x.out", cp.psi.V_x, NX);
        dump("psi_V_z.out", cp.psi.V_z, NZ);
    } else {
        step_velocity(fld, mat, cp);
        dump("V.out", fld.V, N);
        dump("Tz.out", fld.Tz, N);
        dump("psi_Tx_x.out", cp.psi.Tx_x, NX);
        dump("psi_Tz_z.out", cp.psi.Tz_z, NZ);
    }
#define D(arr, comp, n) dump(#arr "_" #comp ".f32", cp.arr.comp, n)
    D(b, Tx_x, nx); D(c, Tx_x, nx); D(k, Tx_x, nx);
    D(b, V_x, nx); D(c, V_x, nx); D(k, V_x, nx);
    D(b, Tz_z, nz); D(c, Tz_z, nz); D(k, Tz_z, nz);
    D(b, V_z, nz); D(c, V_z, nz); D(k, V_z, nz);
    return 0;
}
