// Runs the reference's own SH host precompute (SH/: MATERIAL::get_model,
// CPML::CPML, gaussian/trapozoid) and dumps the arrays.  The PARAM fields
// are filled from key=value arguments here (SH's getpar.cpp does not build
// as C++; par-file parsing is pinned through the P-SV harness, whose
// getpar.c is the same code).  Test infrastructure.
#include "harness_common.h"
#include "param/param.h"
#include "material/material.h"
#include "cpml/cpml.h"
#include "point_source/waveform.h"
#include <cstring>
#include <map>
#include <string>

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
    param.h = Fl("h");
    param.dt = Fl("dt");
    param.npml = I("npml");
    param.pml_dt = Fl("pml_dt");
    param.pml_r = Fl("pml_r");
    param.pml_v = Fl("pml_v");
    param.pml_fc = Fl("pml_fc");
    strcpy(param.modelname, kv.at("model").c_str());
    param.output_material = 0;
    MATERIAL mat;
    mat.init_for_full(param);
    mat.get_model(param);
    const long N = (long)param.nx * param.nz;
    dump("BV.f32", mat.BV, N);
    dump("MUX.f32", mat.MUX, N);
    dump("MUZ.f32", mat.MUZ, N);
    CPML cp(param);
#define D(arr, comp, n) dump(#arr "_" #comp ".f32", cp.arr.comp, n)
    D(b, Tx_x, param.nx); D(c, Tx_x, param.nx); D(k, Tx_x, param.nx);
    D(b, V_x, param.nx); D(c, V_x, param.nx); D(k, V_x, param.nx);
    D(b, Tz_z, param.nz); D(c, Tz_z, param.nz); D(k, Tz_z, param.nz);
    D(b, V_z, param.nz); D(c, V_z, param.nz); D(k, V_z, param.nz);
    float *src = nullptr;
    int lsrc = 0;
    gaussian(param.dt, (float)atof(kv.at("alpha").c_str()), lsrc, src);
    dump("src.f32", src, lsrc);
    return 0;
}
