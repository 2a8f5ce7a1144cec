// Runs the reference's own P-SV FP32 host precompute (PSV/clean_nobox:
// PARAM::read_param, MATERIAL::get_model, CPML::CPML, gaussian/trapozoid)
// on a par file and dumps the arrays it produces.  Test infrastructure.
#include "harness_common.h"
#include "param/param.h"
#include "material/material.h"
#include "cpml/cpml.h"
#include "source/waveform.h"

int main(int ac, char **av) {
    init_outdir();
    PARAM param;
    param.read_param(ac, av, "full");
    MATERIAL mat;
    mat.init_for_full(param);
    mat.get_model(param);
    const long N = (long)param.nx * param.nz;
    dump("BU.f32", mat.BU, N);
    dump("BW.f32", mat.BW, N);
    dump("MU.f32", mat.MU, N);
    dump("MUA.f32", mat.MUA, N);
    dump("LAM.f32", mat.LAM, N);
    CPML cp(param);
#define D(arr, comp, n) dump(#arr "_" #comp ".f32", cp.arr.comp, n)
    D(b, Txx_x, param.nx); D(c, Txx_x, param.nx); D(k, Txx_x, param.nx);
    D(b, Txz_x, param.nx); D(c, Txz_x, param.nx); D(k, Txz_x, param.nx);
    D(b, U_x, param.nx); D(c, U_x, param.nx); D(k, U_x, param.nx);
    D(b, W_x, param.nx); D(c, W_x, param.nx); D(k, W_x, param.nx);
    D(b, Tzz_z, param.nz); D(c, Tzz_z, param.nz); D(k, Tzz_z, param.nz);
    D(b, Txz_z, param.nz); D(c, Txz_z, param.nz); D(k, Txz_z, param.nz);
    D(b, U_z, param.nz); D(c, U_z, param.nz); D(k, U_z, param.nz);
    D(b, W_z, param.nz); D(c, W_z, param.nz); D(k, W_z, param.nz);
    float *src = nullptr;
    int lsrc = 0;
    if (param.srctime[0] == 'T' || param.srctime[0] == 't')
        trapozoid(param.dt, param.trap1, param.trap2, param.trap3, lsrc, src);
    else
        gaussian(param.dt, param.alpha, lsrc, src);
    dump("src.f32", src, lsrc);
    int ints[8] = {param.nx, param.nz, param.npml, param.xs, param.zs, param.nt, param.itrecord, lsrc};
    dump("ints.i32", ints, 8);
    float flts[8] = {param.h, param.dt, param.strike, param.dip, param.rake, param.azimuth, param.pml_r,
                     param.pml_v};
    dump("floats.f32", flts, 8);
    return 0;
}
