/*
 * Oracle template body: included twice by fd_oracle.c, once with
 * REAL=float / SFX=_f32 (the FP32 trees) and once with REAL=double /
 * SFX=_f64 (PSV/psv_clean_double).  TEST INFRASTRUCTURE ONLY: nothing in
 * the product links or calls this code.
 *
 * Every loop below restates the semantics of the reference's GPU kernels
 * (not its CPU fd.cpp loops, which differ at the edges):
 *   - order reduction ord = min(ix, iz, nx-1-ix, nz-1-iz, 4) applied to
 *     both derivatives            (PSV/clean_nobox/fd/cu_fd.cu:165-169)
 *   - any field index outside [0,nx)x[0,nz) reads as zero, which is what the
 *     zero-filled LDS halo produces (cu_fd.cu:72-91)
 *   - the x-branch / z-branch split with free surface at iz=0
 *                                  (cu_fd.cu:190-242, 342-386)
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define F(name) CAT(name, SFX)

/* 8th-order staggered coefficients by order, PSV/clean_nobox/param/const.h:4-18 */
static REAL F(coef_)[5][4];
static int F(coef_ready_) = 0;

static void F(init_coef_)(void)
{
    if (F(coef_ready_)) return;
    const double tab[5][4] = {
        {1.0, 0.0, 0.0, 0.0},
        {1.0, 0.0, 0.0, 0.0},
        {9. / 8., 1. / 24., 0.0, 0.0},
        {75. / 64., 25. / 384., 3. / 640., 0.0},
        {1225. / 1024., 245. / 3072., 49. / 5120., 5. / 7168.}};
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 4; j++) F(coef_)[i][j] = (REAL)tab[i][j];
    F(coef_ready_) = 1;
}

static inline int F(ordof_)(int ix, int iz, int nx, int nz)
{
    int o = ix < iz ? ix : iz;
    if (nx - 1 - ix < o) o = nx - 1 - ix;
    if (nz - 1 - iz < o) o = nz - 1 - iz;
    return o < 4 ? o : 4;
}

/* zero outside the domain (cu_fd.cu:72-91) */
static inline REAL F(at_)(const REAL *f, int nx, int nz, int x, int z)
{
    if (x < 0 || x >= nx || z < 0 || z >= nz) return (REAL)0;
    return f[(size_t)z * nx + x];
}

/* "backward" difference: centre at (x), taps x..x-1 (UDIFF ux, DXDZTXZ dxTxz) */
static inline REAL F(dxm_)(const REAL *f, int nx, int nz, int x, int z, const REAL *c)
{
    return c[0] * (F(at_)(f, nx, nz, x, z) - F(at_)(f, nx, nz, x - 1, z))
         - c[1] * (F(at_)(f, nx, nz, x + 1, z) - F(at_)(f, nx, nz, x - 2, z))
         + c[2] * (F(at_)(f, nx, nz, x + 2, z) - F(at_)(f, nx, nz, x - 3, z))
         - c[3] * (F(at_)(f, nx, nz, x + 3, z) - F(at_)(f, nx, nz, x - 4, z));
}
/* "forward" difference: taps x+1..x (WDIFF wx, DXTXX) */
static inline REAL F(dxp_)(const REAL *f, int nx, int nz, int x, int z, const REAL *c)
{
    return c[0] * (F(at_)(f, nx, nz, x + 1, z) - F(at_)(f, nx, nz, x, z))
         - c[1] * (F(at_)(f, nx, nz, x + 2, z) - F(at_)(f, nx, nz, x - 1, z))
         + c[2] * (F(at_)(f, nx, nz, x + 3, z) - F(at_)(f, nx, nz, x - 2, z))
         - c[3] * (F(at_)(f, nx, nz, x + 4, z) - F(at_)(f, nx, nz, x - 3, z));
}
static inline REAL F(dzm_)(const REAL *f, int nx, int nz, int x, int z, const REAL *c)
{
    return c[0] * (F(at_)(f, nx, nz, x, z) - F(at_)(f, nx, nz, x, z - 1))
         - c[1] * (F(at_)(f, nx, nz, x, z + 1) - F(at_)(f, nx, nz, x, z - 2))
         + c[2] * (F(at_)(f, nx, nz, x, z + 2) - F(at_)(f, nx, nz, x, z - 3))
         - c[3] * (F(at_)(f, nx, nz, x, z + 3) - F(at_)(f, nx, nz, x, z - 4));
}
static inline REAL F(dzp_)(const REAL *f, int nx, int nz, int x, int z, const REAL *c)
{
    return c[0] * (F(at_)(f, nx, nz, x, z + 1) - F(at_)(f, nx, nz, x, z))
         - c[1] * (F(at_)(f, nx, nz, x, z + 2) - F(at_)(f, nx, nz, x, z - 1))
         + c[2] * (F(at_)(f, nx, nz, x, z + 3) - F(at_)(f, nx, nz, x, z - 2))
         - c[3] * (F(at_)(f, nx, nz, x, z + 4) - F(at_)(f, nx, nz, x, z - 3));
}


/* Same expressions with no bounds test, for cells whose 8 taps all lie in
 * the domain (4 <= ix < nx-4 and 4 <= iz < nz-4). */
static inline int F(inner_)(int ix, int iz, int nx, int nz)
{
    return ix >= 4 && ix < nx - 4 && iz >= 4 && iz < nz - 4;
}
#define OR_DXM(f, q, s, c) (c[0] * (f[q] - f[q - s]) - c[1] * (f[q + s] - f[q - 2 * s]) + c[2] * (f[q + 2 * s] - f[q - 3 * s]) - c[3] * (f[q + 3 * s] - f[q - 4 * s]))
#define OR_DXP(f, q, s, c) (c[0] * (f[q + s] - f[q]) - c[1] * (f[q + 2 * s] - f[q - s]) + c[2] * (f[q + 3 * s] - f[q - 2 * s]) - c[3] * (f[q + 4 * s] - f[q - 3 * s]))

/* ------------------------------------------------------------------------
 * C-PML 1-D profile: PSV/clean_nobox/cpml/cpml.cpp:7-83 (FP32 tree keeps
 * the intermediates in float arrays) and PSV/psv_clean_double/cpml/cpml.cpp
 * (all double).  k is stored as 1/kappa; b=c=k=0 for i<4 and i>=n-5.
 * ---------------------------------------------------------------------- */
void F(or_cpml_coef)(int n, int npml, double pml_pos, double startpos, double dt,
                     double R, double v, double fc, REAL *b, REAL *c, REAL *k)
{
    REAL *d = (REAL *)calloc((size_t)n, sizeof(REAL));
    const double pk = 2.0, psig = 2.0, palpha = 1.0;
    const REAL ppos = (REAL)pml_pos;
    const REAL sigmax = (REAL)(-(psig + 1.0) / (2.0 * floor((double)ppos)) * (REAL)v * log((double)(REAL)R));
    const REAL alphamax = (REAL)(3.141592653589 * (REAL)fc);
    const double kmax = 2.0;
    for (int i = 0; i < npml + 2 && i < n; i++) {
        REAL tmp = (REAL)((ppos - (i + (REAL)startpos)) / floor((double)npml));
        if (tmp > 0) d[i] = tmp;
    }
    for (int i = n - npml - 2; i < n; i++) {
        if (i < 0) continue;
        REAL tmp = (REAL)(((i + (REAL)startpos) - (n - 1 - ppos)) / floor((double)npml));
        if (tmp > 0) d[i] = tmp;
    }
    for (int i = 0; i < n; i++) {
        REAL sg = (REAL)(sigmax * pow((double)d[i], psig));
        REAL kp = (REAL)(1.0 + (kmax - 1.0) * pow((double)d[i], pk));
        REAL al = (REAL)(alphamax * (1.0 - pow((double)d[i], palpha)));
        REAL bb = (REAL)exp(-((double)(REAL)(sg / kp) + al) * (double)(REAL)dt);
        REAL cc;
        if ((double)sg + (double)kp * al < 1E-15)
            cc = 0;
        else
            cc = (REAL)(sg * (bb - 1.0) / ((REAL)(sg + kp * al)) / kp);
        b[i] = bb;
        c[i] = cc;
        k[i] = (REAL)(1.0 / kp);
    }
    for (int i = 0; i < 4 && i < n; i++) b[i] = c[i] = k[i] = 0;
    for (int i = n - 5; i < n; i++)
        if (i >= 0) b[i] = c[i] = k[i] = 0;
    free(d);
}

/* ------------------------------------------------------------------------
 * P-SV material: PSV/clean_nobox/material/material.cpp:84-235
 * (FP64: PSV/psv_clean_double/material/material.cpp:101-311).
 * Returns the stability number vpmax*dt*sum(c)*sqrt(2)/h (reference exits
 * when it is > 1).
 * ---------------------------------------------------------------------- */
double F(or_psv_material)(int nx, int nz, double h, double dt, const REAL *vp,
                          const REAL *vs, const REAL *den, REAL *BU, REAL *BW,
                          REAL *MU, REAL *MUA, REAL *LAM)
{
    const size_t N = (size_t)nx * nz;
    const REAL fac = (REAL)((REAL)dt / (REAL)h);
    for (size_t q = 0; q < N; q++) BU[q] = BW[q] = MU[q] = MUA[q] = LAM[q] = 0;
    for (int iz = 0; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++) {
            size_t k = (size_t)iz * nx + ix;
            REAL mu = vs[k] * vs[k] * den[k];
            REAL lam = (REAL)(vp[k] * vp[k] * den[k] - 2.0 * mu);
            LAM[k] = lam;
            MU[k] = mu;
        }
    /* water stabiliser, 5-cell bands on the left, right and bottom (:94-129) */
    const int wst = 5;
#define OR_WATER(k_)                                                     \
    do {                                                                 \
        size_t kk = (k_);                                                \
        if (vs[kk] / vp[kk] < 0.01) {                                    \
            REAL mu = (REAL)(vp[kk] * vp[kk] / 3.0 * den[kk]);           \
            REAL lam = (REAL)(vp[kk] * vp[kk] * den[kk] - 2.0 * mu);     \
            LAM[kk] = lam;                                               \
            MU[kk] = mu;                                                 \
        }                                                                \
    } while (0)
    for (int iz = 0; iz < nz; iz++) {
        for (int ix = 0; ix < wst && ix < nx; ix++) OR_WATER((size_t)iz * nx + ix);
        for (int ix = nx - wst; ix < nx; ix++)
            if (ix >= 0) OR_WATER((size_t)iz * nx + ix);
    }
    for (int iz = nz - wst; iz < nz; iz++) {
        if (iz < 0) continue;
        for (int ix = 0; ix < nx; ix++) OR_WATER((size_t)iz * nx + ix);
    }
#undef OR_WATER
    /* staggered averages (:132-161), including the MUA denominator quirk */
    for (int iz = 1; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++) {
            size_t k = (size_t)iz * nx + ix;
            if (k + 1 < N)
                BU[k] = (REAL)(2.0 / ((double)den[k] + den[k + 1]));
            else
                BU[k] = (REAL)(2.0 / ((double)den[k] + den[k]));
            if (k + nx < N)
                BW[k] = (REAL)(2.0 / ((double)den[k] + den[k - nx]));
            else
                BW[k] = (REAL)(2.0 / ((double)den[k] + den[k]));
            if (ix + 1 < nx && iz + 1 < nz) {
                double m1 = MU[k], m2 = MU[k + 1], m3 = MU[k - nx], m4 = MU[k + 1 - nx];
                double den4 = m2 * m3 * m4 + m1 * m3 * m4 + m1 * m2 * m4 + m2 * m3 * m4;
                if (den4 < 1E-10) den4 = 1E-10;
                MUA[k] = (REAL)(4.0 * m1 * m2 * m3 * m4 / den4);
            } else {
                MUA[k] = MU[k];
            }
        }
    /* Mittet free surface on row 0, ix in [1, nx-2] (:188-201) */
    for (int ix = 1; ix < nx - 1; ix++) {
        if (ix + 1 < nx)
            BU[ix] = (REAL)(4.0 / ((double)den[ix] + den[ix + 1]));
        else
            BU[ix] = (REAL)(4.0 / ((double)den[ix] + den[ix]));
        LAM[ix] = 0;
        MU[ix] = (REAL)(0.5 * MU[ix]);
    }
    for (size_t q = 0; q < N; q++) {
        MU[q] *= fac;
        LAM[q] *= fac;
        BU[q] *= fac;
        BW[q] *= fac;
        MUA[q] *= fac;
    }
    double vpmax = -99999.0;
    for (size_t q = 0; q < N; q++)
        if (vp[q] > vpmax) vpmax = vp[q];
    const double coefsum = 1225. / 1024. + 245. / 3072. + 49. / 5120. + 5. / 7168.;
    return vpmax * (REAL)dt * coefsum * sqrt(2.0) / (REAL)h;
}

/* SH material: SH/material/material.cpp:76-130 */
double F(or_sh_material)(int nx, int nz, double h, double dt, const REAL *vs,
                         const REAL *den, REAL *BV, REAL *MUX, REAL *MUZ)
{
    const size_t N = (size_t)nx * nz;
    const REAL fac = (REAL)((REAL)dt / (REAL)h);
    for (int iz = 0; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++) {
            size_t k = (size_t)iz * nx + ix;
            BV[k] = (REAL)(1.0 / den[k]);
            double m1 = vs[k] * vs[k] * den[k], m2;
            if (k + 1 < N)
                m2 = vs[k + 1] * vs[k + 1] * den[k + 1];
            else
                m2 = m1;
            double s = m1 + m2;
            if (s < 1E-10) s = 1E-10;
            MUX[k] = (REAL)(2.0 * m1 * m2 / s);
            m1 = vs[k] * vs[k] * den[k];
            if ((long)k - nx >= 0)
                m2 = vs[k - nx] * vs[k - nx] * den[k - nx];
            else
                m2 = m1;
            s = m1 + m2;
            if (s < 1E-10) s = 1E-10;
            MUZ[k] = (REAL)(2.0 * m1 * m2 / s);
        }
    for (size_t q = 0; q < N; q++) {
        MUX[q] *= fac;
        MUZ[q] *= fac;
        BV[q] *= fac;
    }
    double vsmax = -99999.0;
    for (size_t q = 0; q < N; q++)
        if (vs[q] > vsmax) vsmax = vs[q];
    const double coefsum = 1225. / 1024. + 245. / 3072. + 49. / 5120. + 5. / 7168.;
    return vsmax * (REAL)dt * coefsum * sqrt(2.0) / (REAL)h;
}

/* ------------------------------------------------------------------------
 * P-SV stress pass: PSV/clean_nobox/fd/cu_fd.cu:96-245.
 * psi/b/c/k order: [0]=U_x [1]=W_x [2]=U_z [3]=W_z.
 * ---------------------------------------------------------------------- */
void F(or_psv_stress)(int nx, int nz, int npml, int topabsorb, const REAL *U,
                      const REAL *W, REAL *Txx, REAL *Tzz, REAL *Txz,
                      const REAL *LAM, const REAL *MU, const REAL *MUA,
                      REAL *const *psi, const REAL *const *b,
                      const REAL *const *c, const REAL *const *k,
                      const unsigned char *open)
{
    F(init_coef_)();
    REAL *pUx = psi[0], *pWx = psi[1], *pUz = psi[2], *pWz = psi[3];
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++) {
            const size_t in = (size_t)iz * nx + ix;
            if (open && !open[in]) continue; /* fd/cu_fd.cu:16-23 */
            const int ord = F(ordof_)(ix, iz, nx, nz);
            const REAL *cf = F(coef_)[ord];
            const REAL lam = LAM[in], mu = MU[in], mua = MUA[in];
            const REAL gam = lam + (REAL)2.0 * mu;
            const int inr = F(inner_)(ix, iz, nx, nz);
            const REAL ux = inr ? OR_DXM(U, in, 1, cf) : F(dxm_)(U, nx, nz, ix, iz, cf);
            const REAL uz = inr ? OR_DXM(U, in, (size_t)nx, cf) : F(dzm_)(U, nx, nz, ix, iz, cf);
            const REAL wx = inr ? OR_DXP(W, in, 1, cf) : F(dxp_)(W, nx, nz, ix, iz, cf);
            const REAL wz = inr ? OR_DXP(W, in, (size_t)nx, cf) : F(dzp_)(W, nx, nz, ix, iz, cf);
            size_t idx;
            if (ix < npml || ix >= nx - npml) {
                idx = (ix < npml) ? (size_t)ix + (size_t)iz * 2 * npml
                                  : (size_t)(npml + nx - 1 - ix) + (size_t)iz * 2 * npml;
                pUx[idx] = b[0][ix] * pUx[idx] + c[0][ix] * ux;
                Txx[in] += gam * (ux * k[0][ix] + pUx[idx]);
                Tzz[in] += lam * (ux * k[0][ix] + pUx[idx]);
                pWx[idx] = b[1][ix] * pWx[idx] + c[1][ix] * wx;
                Txz[in] += mua * (wx * k[1][ix] + pWx[idx]);
            } else if (iz != 0) {
                Txx[in] += gam * (ux);
                Tzz[in] += lam * (ux);
                Txz[in] += mua * (wx);
            }
            if ((iz < npml && topabsorb) || iz >= nz - npml) {
                idx = (iz < npml && topabsorb) ? (size_t)iz * nx + ix
                                               : (size_t)(npml + nz - 1 - iz) * nx + ix;
                pWz[idx] = b[3][iz] * pWz[idx] + c[3][iz] * wz;
                Txx[in] += lam * (wz * k[3][iz] + pWz[idx]);
                Tzz[in] += gam * (wz * k[3][iz] + pWz[idx]);
                pUz[idx] = b[2][iz] * pUz[idx] + c[2][iz] * uz;
                Txz[in] += mua * (uz * k[2][iz] + pUz[idx]);
            } else if (iz == 0) {
                Txx[in] += (REAL)2.0 * mu * ux;
                Tzz[in] = 0;
                Txz[in] = 0;
            } else {
                Txx[in] += lam * (wz);
                Tzz[in] += gam * (wz);
                Txz[in] += mua * (uz);
            }
        }
}

/* P-SV velocity pass: PSV/clean_nobox/fd/cu_fd.cu:246-388.
 * psi/b/c/k order: [0]=Txx_x [1]=Txz_x [2]=Txz_z [3]=Tzz_z. */
void F(or_psv_velocity)(int nx, int nz, int npml, int topabsorb, const REAL *Txx,
                        const REAL *Tzz, const REAL *Txz, REAL *U, REAL *W,
                        const REAL *BU, const REAL *BW, REAL *const *psi,
                        const REAL *const *b, const REAL *const *c,
                        const REAL *const *k, const unsigned char *open)
{
    F(init_coef_)();
    REAL *pTxx_x = psi[0], *pTxz_x = psi[1], *pTxz_z = psi[2], *pTzz_z = psi[3];
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++) {
            const size_t in = (size_t)iz * nx + ix;
            if (open && !open[in]) continue;
            const int ord = F(ordof_)(ix, iz, nx, nz);
            const REAL *cf = F(coef_)[ord];
            const REAL bu = BU[in], bw = BW[in];
            const int inr = F(inner_)(ix, iz, nx, nz);
            const REAL dxTxx = inr ? OR_DXP(Txx, in, 1, cf) : F(dxp_)(Txx, nx, nz, ix, iz, cf);
            const REAL dzTzz = inr ? OR_DXM(Tzz, in, (size_t)nx, cf) : F(dzm_)(Tzz, nx, nz, ix, iz, cf);
            const REAL dxTxz = inr ? OR_DXM(Txz, in, 1, cf) : F(dxm_)(Txz, nx, nz, ix, iz, cf);
            const REAL dzTxz = inr ? OR_DXP(Txz, in, (size_t)nx, cf) : F(dzp_)(Txz, nx, nz, ix, iz, cf);
            size_t idx;
            if (ix < npml || ix >= nx - npml) {
                idx = (ix < npml) ? (size_t)ix + (size_t)iz * 2 * npml
                                  : (size_t)(npml + nx - 1 - ix) + (size_t)iz * 2 * npml;
                pTxx_x[idx] = b[0][ix] * pTxx_x[idx] + c[0][ix] * dxTxx;
                U[in] += bu * (dxTxx * k[0][ix] + pTxx_x[idx]);
                pTxz_x[idx] = b[1][ix] * pTxz_x[idx] + c[1][ix] * dxTxz;
                W[in] += bw * (dxTxz * k[1][ix] + pTxz_x[idx]);
            } else if (iz != 0) {
                U[in] += bu * (dxTxx);
                W[in] += bw * (dxTxz);
            }
            if ((iz < npml && topabsorb) || iz >= nz - npml) {
                idx = (iz < npml && topabsorb) ? (size_t)iz * nx + ix
                                               : (size_t)(npml + nz - 1 - iz) * nx + ix;
                pTxz_z[idx] = b[2][iz] * pTxz_z[idx] + c[2][iz] * dzTxz;
                pTzz_z[idx] = b[3][iz] * pTzz_z[idx] + c[3][iz] * dzTzz;
                U[in] += bu * (dzTxz * k[2][iz] + pTxz_z[idx]);
                W[in] += bw * (dzTzz * k[3][iz] + pTzz_z[idx]);
            } else if (iz == 0) {
                U[in] += bu * (dxTxx + dzTxz);
                W[in] = 0;
            } else {
                U[in] += bu * (dzTxz);
                W[in] += bw * (dzTzz);
            }
        }
}

/* SH stress pass: SH/fd/cu_fd.cu:77-171.  psi order [0]=V_x [1]=V_z. */
void F(or_sh_stress)(int nx, int nz, int npml, int topabsorb, const REAL *V, REAL *Tx,
                     REAL *Tz, const REAL *MUX, const REAL *MUZ, REAL *const *psi,
                     const REAL *const *b, const REAL *const *c, const REAL *const *k,
                     const unsigned char *open)
{
    F(init_coef_)();
    REAL *pVx = psi[0], *pVz = psi[1];
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++) {
            if (open && !open[(size_t)iz * nx + ix]) continue; /* SH/fd/cu_fd.cu:106 */
            /* early return after the derivatives (:140) */
            if (ix == 0 || ix >= nx - 1 || iz >= nz - 1) continue;
            const size_t in = (size_t)iz * nx + ix;
            const int ord = F(ordof_)(ix, iz, nx, nz);
            const REAL *cf = F(coef_)[ord];
            const REAL mux = MUX[in], muz = MUZ[in];
            const REAL dxV = F(dxp_)(V, nx, nz, ix, iz, cf);
            const REAL dzV = F(dzm_)(V, nx, nz, ix, iz, cf);
            size_t idx;
            if (ix < npml || ix >= nx - npml) {
                idx = (ix < npml) ? (size_t)ix + (size_t)iz * 2 * npml
                                  : (size_t)(npml + nx - 1 - ix) + (size_t)iz * 2 * npml;
                pVx[idx] = b[0][ix] * pVx[idx] + c[0][ix] * dxV;
                Tx[in] += mux * (dxV * k[0][ix] + pVx[idx]);
            } else {
                Tx[in] += mux * (dxV);
            }
            if ((iz < npml && topabsorb) || iz >= nz - npml) {
                idx = (iz < npml && topabsorb) ? (size_t)iz * nx + ix
                                               : (size_t)(npml + nz - 1 - iz) * nx + ix;
                pVz[idx] = b[1][iz] * pVz[idx] + c[1][iz] * dzV;
                Tz[in] += muz * (dzV * k[1][iz] + pVz[idx]);
            } else {
                Tz[in] += muz * (dzV);
            }
        }
}

/* SH velocity pass: SH/fd/cu_fd.cu:172-273, Tz row-0 imaging first (:226).
 * psi order [0]=Tx_x [1]=Tz_z. */
void F(or_sh_velocity)(int nx, int nz, int npml, int topabsorb, REAL *Tx, REAL *Tz,
                       REAL *V, const REAL *BV, REAL *const *psi,
                       const REAL *const *b, const REAL *const *c, const REAL *const *k,
                       const unsigned char *open)
{
    F(init_coef_)();
    REAL *pTx = psi[0], *pTz = psi[1];
    /* the imaging runs after the window check (:206, :226) */
    if (nz > 1)
        for (int ix = 0; ix < nx; ix++)
            if (!open || open[ix]) Tz[ix] = -Tz[ix + nx];
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++) {
            const size_t in = (size_t)iz * nx + ix;
            if (open && !open[in]) continue;
            const int ord = F(ordof_)(ix, iz, nx, nz);
            const REAL *cf = F(coef_)[ord];
            const REAL bv = BV[in];
            const REAL Tx_x = F(dxm_)(Tx, nx, nz, ix, iz, cf);
            const REAL Tz_z = F(dzp_)(Tz, nx, nz, ix, iz, cf);
            size_t idx;
            if (ix < npml || ix >= nx - npml) {
                idx = (ix < npml) ? (size_t)ix + (size_t)iz * 2 * npml
                                  : (size_t)(npml + nx - 1 - ix) + (size_t)iz * 2 * npml;
                pTx[idx] = b[0][ix] * pTx[idx] + c[0][ix] * Tx_x;
                V[in] += bv * (Tx_x * k[0][ix] + pTx[idx]);
            } else {
                V[in] += bv * (Tx_x);
            }
            if ((iz < npml && topabsorb) || iz >= nz - npml) {
                idx = (iz < npml && topabsorb) ? (size_t)iz * nx + ix
                                               : (size_t)(npml + nz - 1 - iz) * nx + ix;
                pTz[idx] = b[1][iz] * pTz[idx] + c[1][iz] * Tz_z;
                V[in] += bv * (Tz_z * k[1][iz] + pTz[idx]);
            } else {
                V[in] += bv * (Tz_z);
            }
        }
}

/* ------------------------------------------------------------------------
 * Residual box grid (hybrid coupling): PSV/psv_clean_double/fd/cu_fd.cu:93-207.
 * Full 8th order everywhere, zero outside the bx*bz grid, no PML.
 * ---------------------------------------------------------------------- */
void F(or_box_stress)(int bx, int bz, const REAL *U, const REAL *W, REAL *Txx, REAL *Tzz,
                      REAL *Txz, const REAL *LAM, const REAL *MU, const REAL *MUA)
{
    F(init_coef_)();
    const REAL *cf = F(coef_)[4];
    for (int iz = 0; iz < bz; iz++)
        for (int ix = 0; ix < bx; ix++) {
            const size_t in = (size_t)iz * bx + ix;
            const REAL lam = LAM[in], mua = MUA[in];
            const REAL gam = lam + (REAL)2.0 * MU[in];
            const REAL ux = F(dxm_)(U, bx, bz, ix, iz, cf);
            const REAL uz = F(dzm_)(U, bx, bz, ix, iz, cf);
            const REAL wx = F(dxp_)(W, bx, bz, ix, iz, cf);
            const REAL wz = F(dzp_)(W, bx, bz, ix, iz, cf);
            Txx[in] += gam * ux + lam * wz;
            Tzz[in] += lam * ux + gam * wz;
            Txz[in] += mua * (wx + uz);
        }
}

void F(or_box_velocity)(int bx, int bz, const REAL *Txx, const REAL *Tzz, const REAL *Txz,
                        REAL *U, REAL *W, const REAL *BU, const REAL *BW)
{
    F(init_coef_)();
    const REAL *cf = F(coef_)[4];
    for (int iz = 0; iz < bz; iz++)
        for (int ix = 0; ix < bx; ix++) {
            const size_t in = (size_t)iz * bx + ix;
            const REAL dxTxx = F(dxp_)(Txx, bx, bz, ix, iz, cf);
            const REAL dzTzz = F(dzm_)(Tzz, bx, bz, ix, iz, cf);
            const REAL dxTxz = F(dxm_)(Txz, bx, bz, ix, iz, cf);
            const REAL dzTxz = F(dzp_)(Txz, bx, bz, ix, iz, cf);
            U[in] += BU[in] * (dxTxx + dzTxz);
            W[in] += BW[in] * (dxTxz + dzTzz);
        }
}

/* Ring coupling, PSV/psv_clean_double/box/box.cpp:198-210:
 * sub ring r[ri] = a[ai] - s[ri]; add ring a[ai] = r[ri] + s[ri]. */
void F(or_ring_sub)(int n, const int *aidx, const int *ridx, const REAL *a, REAL *r,
                    const REAL *s)
{
    for (int i = 0; i < n; i++) r[ridx[i]] = a[aidx[i]] - s[ridx[i]];
}
void F(or_ring_add)(int n, const int *aidx, const int *ridx, REAL *a, const REAL *r,
                    const REAL *s)
{
    for (int i = 0; i < n; i++) a[aidx[i]] = r[ridx[i]] + s[ridx[i]];
}

/* Receiver gather, SH/record/cu_record.cu:136-144 (flat index rx + rz*nx,
 * out-of-array reads return zero like the zeroed device halo rows). */
void F(or_record)(int nx, int nz, const REAL *f, int n, const int *rx, const int *rz,
                  REAL *out)
{
    const long N = (long)nx * nz;
    for (int i = 0; i < n; i++) {
        long q = (long)rx[i] + (long)rz[i] * nx;
        out[i] = (q >= 0 && q < N) ? f[q] : (REAL)0;
    }
}

/* Read-modify-write of one cell, as the reference's host round trip does
 * (PSV/clean_nobox/point_source/cu_point_source.cu:47-51). */
void F(or_add_cell)(REAL *f, int nx, int nz, int x, int z, double v)
{
    if (x < 0 || x >= nx || z < 0 || z >= nz) return;
    REAL t = f[(size_t)z * nx + x];
    t = (REAL)(t + (REAL)v);
    f[(size_t)z * nx + x] = t;
}

/* ------------------------------------------------------------------------
 * Point sources.  kind: 0 = line moment source ('1'), 1 = cumulative
 * moment source ('p'), 2 = explosion (-DEXPLOSION_SOURCE build of '1'),
 * 3 = explosion with the 'p' injection pattern.
 * Moment tensor and scaling: PSV/clean_nobox/point_source/cu_point_source.cu:16-40,
 * injection pattern :45-70 (line) and :110-158 ('p').
 * ---------------------------------------------------------------------- */
void F(or_psv_moment)(int kind, double strike, double dip, double rake, double azimuth,
                      double h, double dt, REAL *M /* Mxx, Mxz, Mzz */)
{
    const double PI_ = 3.14159265358979;
    REAL Mxx, Mxz, Mzz;
    if (kind == 2 || kind == 3) {
        Mxx = 1; Mzz = 1; Mxz = 0;
    } else {
        REAL delta = (REAL)((REAL)dip * PI_ / 180.0);
        REAL lambda = (REAL)((REAL)rake * PI_ / 180.0);
        REAL phi = (REAL)(((REAL)strike - (REAL)azimuth) * PI_ / 180.0);
        Mxx = (REAL)(-(sin(delta) * cos(lambda) * sin(2.0 * phi) +
                       sin(2.0 * delta) * sin(lambda) * sin(phi) * sin(phi)));
        Mxz = (REAL)(-(cos(delta) * cos(lambda) * cos(phi) +
                       cos(2.0 * delta) * sin(lambda) * sin(phi)));
        Mzz = (REAL)(sin(2.0 * delta) * sin(lambda));
    }
    double hh = (REAL)h;
    REAL factor = (REAL)(1.0 / PI_ / 1.4142135623731 / (double)(REAL)(hh * hh));
    if (kind == 1 || kind == 3) factor = (REAL)(1.0 / PI_ / 1.4142135623731 / (double)(REAL)(hh * hh) * (REAL)dt / hh);
    M[0] = Mxx * factor;
    M[1] = Mxz * factor;
    M[2] = Mzz * factor;
}

static inline void F(rmw_)(REAL *f, int nx, int nz, int x, int z, REAL delta)
{
    if (x < 0 || x >= nx || z < 0 || z >= nz) return;
    REAL t = f[(size_t)z * nx + x];
    t += delta;
    f[(size_t)z * nx + x] = t;
}

/* src: the STF for kind 0/2 and its running sum for kind 1/3 (main.cpp:94-99). */
void F(or_psv_inject)(int kind, int it, int lsrc, const REAL *src, const REAL *M, int xs,
                      int zs, int nx, int nz, REAL *Txx, REAL *Tzz, REAL *Txz)
{
    const REAL Mxx = M[0], Mxz = M[1], Mzz = M[2];
    if (kind == 0 || kind == 2) {
        if (it >= lsrc) return;
        const REAL s = src[it];
        F(rmw_)(Txx, nx, nz, xs, zs, -(s * Mxx));
        F(rmw_)(Tzz, nx, nz, xs, zs, -(s * Mzz));
        for (int x = xs - 1; x <= xs; x++)
            for (int z = zs; z <= zs + 1; z++)
                F(rmw_)(Txz, nx, nz, x, z, -(REAL)(s * Mxz * (REAL)0.25));
    } else {
        if (it >= lsrc) it = lsrc - 1;
        const REAL s = src[it];
        F(rmw_)(Txx, nx, nz, xs, zs, -(s * Mxx));
        F(rmw_)(Txx, nx, nz, xs - 1, zs, (s * Mxx));
        F(rmw_)(Tzz, nx, nz, xs, zs, -(s * Mzz));
        F(rmw_)(Tzz, nx, nz, xs - 1, zs, (s * Mzz));
        for (int x = xs - 1; x <= xs; x++)
            for (int z = zs; z <= zs + 1; z++) {
                F(rmw_)(Txz, nx, nz, x, z, -(REAL)(s * Mxz * (REAL)0.25));
                F(rmw_)(Txz, nx, nz, x - 1, z, (REAL)(s * Mxz * (REAL)0.25));
            }
    }
}

/* SH moment tensor, SH/point_source/cu_point_source.cu:16-31 (angles in double). */
void F(or_sh_moment)(int kind, double strike, double dip, double rake, double azimuth,
                     double h, double dt, REAL *M /* Mxy, Mzy */)
{
    const double PI_ = 3.14159265358979;
    double delta = (REAL)dip * PI_ / 180.0;
    double lambda = (REAL)rake * PI_ / 180.0;
    double phi = ((REAL)strike - (REAL)azimuth) * PI_ / 180.0;
    REAL Mxy = (REAL)(sin(delta) * cos(lambda) * cos(2.0 * phi) +
                      0.5 * sin(2.0 * delta) * sin(lambda) * sin(2.0 * phi));
    REAL Mzy = (REAL)(-(cos(delta) * cos(lambda) * sin(phi) -
                        cos(2.0 * delta) * sin(lambda) * cos(phi)));
    double hh = (REAL)h;
    REAL factor = (REAL)(1.0 / PI_ / 1.4142135623731 / (double)(REAL)(hh * hh));
    if (kind == 1) factor = (REAL)(1.0 / PI_ / 1.4142135623731 / (double)(REAL)(hh * hh) * (REAL)dt / hh);
    M[0] = Mxy * factor;
    M[1] = Mzy * factor;
}

static inline void F(rmwd_)(REAL *f, int nx, int nz, int x, int z, double delta)
{
    if (x < 0 || x >= nx || z < 0 || z >= nz) return;
    double t = f[(size_t)z * nx + x];
    f[(size_t)z * nx + x] = (REAL)(t + delta);
}

/* SH sources: kind 0 = stype 3 (:6-66), 1 = stype 13 (:68-160), 2 = point
 * force on V (:162-201).  The 0.5*src*M products are double, as written. */
void F(or_sh_inject)(int kind, int it, int lsrc, const REAL *src, const REAL *M, int xs,
                     int zs, int nx, int nz, REAL *Tx, REAL *Tz, REAL *V)
{
    const double Mxy = M[0], Mzy = M[1];
    if (kind == 1 && it >= lsrc) it = lsrc - 1;
    if (it >= lsrc) return;
    const double s = src[it];
    if (kind == 2) {
        F(rmwd_)(V, nx, nz, xs, zs, -(0.5 * s * Mxy));
        return;
    }
    F(rmwd_)(Tx, nx, nz, xs, zs, -(0.5 * s * Mxy));
    F(rmwd_)(Tx, nx, nz, xs - 1, zs, -(0.5 * s * Mxy));
    F(rmwd_)(Tz, nx, nz, xs, zs, -(0.5 * s * Mzy));
    F(rmwd_)(Tz, nx, nz, xs, zs + 1, -(0.5 * s * Mzy));
    if (kind == 1) {
        F(rmwd_)(Tx, nx, nz, xs - 1, zs, (0.5 * s * Mxy));
        F(rmwd_)(Tx, nx, nz, xs - 2, zs, (0.5 * s * Mxy));
        F(rmwd_)(Tz, nx, nz, xs - 1, zs, (0.5 * s * Mzy));
        F(rmwd_)(Tz, nx, nz, xs - 1, zs + 1, (0.5 * s * Mzy));
    }
}

/* ------------------------------------------------------------------------
 * Source-time functions, PSV/clean_nobox/source/waveform.cpp:8-47.
 * gaussian: alpha arrives as int (Q14); returns lsrc; out needs 2*|3a|+2.
 * ---------------------------------------------------------------------- */
int F(or_gaussian)(int alpha, REAL *out)
{
    int aint = abs((int)(alpha * 3.0));
    int lsrc = 2 * aint + 1;
    if (alpha >= 0.0) {
        for (int i = -aint; i <= aint; i++)
            out[i + aint] = (REAL)((REAL)i * (REAL)exp((double)((REAL)(-i) * (REAL)i / (REAL)(alpha * alpha))));
    } else {
        REAL sum = 0;
        for (int i = -aint; i <= aint; i++) {
            out[i + aint] = (REAL)exp((double)((REAL)(-i) * (REAL)i / (REAL)(alpha * alpha)));
            sum += out[i + aint];
        }
        for (int i = 0; i < lsrc; i++) out[i] /= sum;
    }
    return lsrc;
}

/* trapozoid (waveform.cpp:8-26): returns lsrc; out needs lsrc+1 entries. */
int F(or_trapezoid)(double dt, double trap1, double trap2, double trap3, REAL *out)
{
    REAL dtr = (REAL)dt, t1r = (REAL)trap1, t2r = (REAL)trap2, t3r = (REAL)trap3;
    int t1 = (int)(t1r / dtr);
    int t2 = (int)(t2r / dtr);
    int t3 = (int)(t3r / dtr);
    int i;
    t2 += t1;
    int lsrc = t2 + t3;
    REAL amp = (REAL)((t2r + 0.5 * (t1r + t3r)) / dtr);
    REAL *p = out;
    for (i = 0; i < t1; i++) *p++ = (REAL)(i / (t1 * amp));
    for (; i < t2; i++) *p++ = (REAL)(1.0 / amp);
    for (i = t3; i > 0; i--) *p++ = (REAL)(i / (t3 * amp));
    return lsrc;
}

#undef OR_DXM
#undef OR_DXP
#undef F
#undef CAT
#undef CAT2
