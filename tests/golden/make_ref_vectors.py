"""Write tests/golden/ref_host/*.npz from the reference's own host code.

Runs oracle/_ref/ref_psv, ref_sh and ref_sh_fd (the reference's material,
C-PML, source-time-function, model and getpar sources and its SH CPU loop,
compiled by oracle/ref/Makefile) on small synthetic inputs and stores inputs and
outputs as fixtures, so the tests pin the host precompute without the
reference tree.  Needs /root/reference; run from the repo root:
    make -C oracle/ref && python tests/golden/make_ref_vectors.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = os.path.join(ROOT, "oracle", "_ref")
OUT = os.path.join(ROOT, "tests", "golden", "ref_host")


def layered(nx, nz, seed, water=False):
    rng = np.random.default_rng(seed)
    vp = np.empty((nz, nx), np.float32)
    vs = np.empty((nz, nx), np.float32)
    den = np.empty((nz, nx), np.float32)
    edges = [0, nz // 4, nz // 2, 3 * nz // 4, nz]
    for i in range(4):
        v = rng.uniform(2.0, 6.0)
        vp[edges[i]:edges[i + 1]] = v
        vs[edges[i]:edges[i + 1]] = v / rng.uniform(1.6, 2.0)
        den[edges[i]:edges[i + 1]] = rng.uniform(2.0, 3.0)
    vp *= 1 + 0.02 * rng.standard_normal((nz, nx)).astype(np.float32)
    vs *= 1 + 0.02 * rng.standard_normal((nz, nx)).astype(np.float32)
    if water:
        # fluid cells in the left/right/bottom stabiliser bands and inside
        vs[:, :3] = 0.0
        vs[-2:, :] = 0.0
        vs[10:14, 20:30] = 0.0
    return vp, vs, den


def read(d, name, dtype, n=None):
    a = np.fromfile(os.path.join(d, name), dtype)
    return a if n is None else a[:n]


def run(exe, args, d):
    env = dict(os.environ, FD23_REF_OUT=d)
    subprocess.run([exe] + args, env=env, check=True, cwd=d, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL)


PSV_CPML = ["Txx_x", "Txz_x", "U_x", "W_x", "Tzz_z", "Txz_z", "U_z", "W_z"]
SH_CPML = ["Tx_x", "V_x", "Tz_z", "V_z"]

PAR_TEXT = """# fixture par file: comments, several entries per line, quoting,
# substitution and a nested par file
nx=64 nz=48   h=0.25 dt=0.01
nt=300   itrecord=2 # trailing comment
model=mod
xs=30 zs=$(nested_zs)
srctime=G alpha=-7.9
strike=10 dip=60 rake="45" azimuth=5
npml=8 pml_dt=0.01 pml_r=1e-6 pml_v=40 pml_fc=3
output=out nrec=10 ixrec0=5 izrec0_u=2 izrec0_w=3 idxrec=4 idzrec=0
ntsnap=1000
par=nested.par
dip=61
"""
NESTED_TEXT = "nested_zs=12\nstrike=11 \\\n  rake=46\n"


def main():
    os.makedirs(OUT, exist_ok=True)
    exe_p, exe_s = os.path.join(REF, "ref_psv"), os.path.join(REF, "ref_sh")
    if not (os.path.exists(exe_p) and os.path.exists(exe_s)):
        sys.exit("build oracle/_ref first: make -C oracle/ref")
    cases = [
        ("psv_layered", 64, 48, 1, False, "G", -7.9),
        ("psv_water", 64, 48, 2, True, "G", 7.9),
        ("psv_trap", 40, 36, 3, False, "T", 0.0),
    ]
    for name, nx, nz, seed, water, st, alpha in cases:
        with tempfile.TemporaryDirectory() as d:
            vp, vs, den = layered(nx, nz, seed, water)
            for a, ext in ((vp, "vp"), (vs, "vs"), (den, "den")):
                a.tofile(os.path.join(d, "mod." + ext))
            text = PAR_TEXT.replace("nx=64 nz=48", "nx=%d nz=%d" % (nx, nz))
            text = text.replace("srctime=G alpha=-7.9", "srctime=%s alpha=%s trap1=0.05 trap2=0.03 trap3=0.08"
                                % (st, alpha))
            open(os.path.join(d, "case.par"), "w").write(text)
            open(os.path.join(d, "nested.par"), "w").write(NESTED_TEXT)
            run(exe_p, ["par=case.par", "pml_fc=2.5"], d)
            N = nx * nz
            out = {"vp": vp, "vs": vs, "den": den, "par": np.frombuffer(text.encode(), np.uint8),
                   "nested": np.frombuffer(NESTED_TEXT.encode(), np.uint8)}
            for k in ("BU", "BW", "MU", "MUA", "LAM"):
                out[k] = read(d, k + ".f32", np.float32).reshape(nz, nx)
            for c in PSV_CPML:
                n = nx if c.endswith("_x") else nz
                for a in "bck":
                    out["%s_%s" % (a, c)] = read(d, "%s_%s.f32" % (a, c), np.float32, n)
            out["src"] = read(d, "src.f32", np.float32)
            out["ints"] = read(d, "ints.i32", np.int32)
            out["floats"] = read(d, "floats.f32", np.float32)
            np.savez_compressed(os.path.join(OUT, name + ".npz"), **out)
            print(name, "ok")
    # a polygon model with DEF symbols, rasterised by get_elas_model
    mdl = ("BACK 5.0 3.0 2.5\nDEF A 2.0 1.0\nDEF B 10.0 3.0\n"
           "POLY 6.0 -0.55 2.7\n\tA B\n\t4.0 9.0\nPOLY 7.0 4.0 3.0\n\t0 8\n\t16 8\n\t16 12\n\t0 12\n")
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "m.mdl"), "w").write(mdl)
        text = PAR_TEXT.replace("model=mod", "model=m.mdl")
        open(os.path.join(d, "case.par"), "w").write(text)
        open(os.path.join(d, "nested.par"), "w").write(NESTED_TEXT)
        run(exe_p, ["par=case.par"], d)
        nx, nz = 64, 48
        out = {"mdl": np.frombuffer(mdl.encode(), np.uint8)}
        for k in ("BU", "BW", "MU", "MUA", "LAM"):
            out[k] = read(d, k + ".f32", np.float32).reshape(nz, nx)
        np.savez_compressed(os.path.join(OUT, "psv_mdl.npz"), **out)
        print("psv_mdl ok")
    # SH material and C-PML
    with tempfile.TemporaryDirectory() as d:
        nx, nz = 60, 44
        vp, vs, den = layered(nx, nz, 5)
        vs.tofile(os.path.join(d, "mod.vs"))
        den.tofile(os.path.join(d, "mod.den"))
        args = ["nx=%d" % nx, "nz=%d" % nz, "h=0.2", "dt=0.01", "npml=7", "pml_dt=0.01", "pml_r=1e-6",
                "pml_v=30", "pml_fc=2", "model=mod", "alpha=-6"]
        run(exe_s, args, d)
        out = {"vs": vs, "den": den}
        for k in ("BV", "MUX", "MUZ"):
            out[k] = read(d, k + ".f32", np.float32).reshape(nz, nx)
        for c in SH_CPML:
            n = nx if c.endswith("_x") else nz
            for a in "bck":
                out["%s_%s" % (a, c)] = read(d, "%s_%s.f32" % (a, c), np.float32, n)
        out["src"] = read(d, "src.f32", np.float32)
        np.savez_compressed(os.path.join(OUT, "sh_layered.npz"), **out)
        print("sh_layered ok")
    sh_fd()


def sh_fd():
    """One stress pass and one velocity pass of the reference's SH CPU loop
    (SH/fd/fd.cpp) from random fields, psi and material, with its C-PML
    profiles (npml 8, bottom and side strips)."""
    exe = os.path.join(REF, "ref_sh_fd")
    nx, nz, npml = 48, 40, 8
    rng = np.random.default_rng(77)
    f32 = lambda *s: rng.standard_normal(s).astype(np.float32)
    ins = {"V": f32(nz, nx), "Tx": f32(nz, nx), "Tz": f32(nz, nx),
           "BV": rng.uniform(0.1, 0.3, (nz, nx)).astype(np.float32),
           "MUX": rng.uniform(0.5, 2.0, (nz, nx)).astype(np.float32),
           "MUZ": rng.uniform(0.5, 2.0, (nz, nx)).astype(np.float32),
           "psi_V_x": f32(2 * npml * nz), "psi_V_z": f32(2 * npml * nx),
           "psi_Tx_x": f32(2 * npml * nz), "psi_Tz_z": f32(2 * npml * nx)}
    args = ["nx=%d" % nx, "nz=%d" % nz, "npml=%d" % npml, "pml_dt=0.01", "pml_r=1e-6", "pml_v=30", "pml_fc=2"]
    out = {"in_" + k: v for k, v in ins.items()}
    for mode, names in (("stress", ("Tx", "Tz", "psi_V_x", "psi_V_z")),
                        ("velocity", ("V", "Tz", "psi_Tx_x", "psi_Tz_z"))):
        with tempfile.TemporaryDirectory() as d:
            for k, v in ins.items():
                v.tofile(os.path.join(d, k + ".in"))
            run(exe, args + ["mode=" + mode], d)
            for k in names:
                out["%s_%s" % (mode, k)] = read(d, k + ".out", np.float32)
            for c in SH_CPML:
                n = nx if c.endswith("_x") else nz
                for a in "bck":
                    out["%s_%s" % (a, c)] = read(d, "%s_%s.f32" % (a, c), np.float32, n)
    out["dims"] = np.array([nx, nz, npml], np.int32)
    np.savez_compressed(os.path.join(OUT, "sh_fd.npz"), **out)
    print("sh_fd ok")


if __name__ == "__main__":
    main()
