/*
 * Body of the isotropic elastic standard-staggered-grid oracle with a free
 * surface, included twice by elastic.c (REAL = float / double, FN(x) names
 * the instance).  TEST INFRASTRUCTURE ONLY.
 *
 * Restates forward/yang_version_2016_2_2 ("pub" =
 * Iso_multishot_fd_subroutine_fsbc_zy_m2_pub.f90, "pri" =
 * Iso_multishot_fd_subroutine_fsbc_zy_m2_pri.f90), free-surface mode:
 *   substagger            pub:323-547   staggered moduli and buoyancies,
 *                                       replicated above the topography
 *   csponge               pub:796-847   cosine sponge tables
 *   calpmlbclrb           pub:895-934   left / right / bottom split sponge
 *   calfsbcv              pub:938-972   vz above the surface from tau_zz = 0
 *   calfsbcstrzz          pub:978-1001  tau_zz = 0 at and above the surface
 *   calfsbcstrxz          pub:1009-1023 tau_xz image above the surface
 *   calux / caluz         pub:1026-1084, 1160-1337 (order reduced near the surface)
 *   calu, addsrc, get_gauss_filter_coe  pub:1412-1501
 *   get_record            pub:2-16
 *   isoewm2dfsbc          pri:57-230   the time loop, in its call order
 * Layout: [nx][nz] (z fastest), every field with H = 5 zero cells on each
 * side (the Fortran -4:nz+5 bounds; split fields 0:nz need row -1).
 * Fortran (k, i) = C (i - 1, k - 1); itopo[] holds 0-based surface rows.
 */

#define H 5
#define NZH (nz + 2 * H)
#define UF(u, i, k) (u)[(size_t)((i) + H) * NZH + (size_t)((k) + H)]
#define AT(u, i, k) (u)[(size_t)(i) * nz + (size_t)(k)]

/* Taylor tables of orders 2..10, the module constants (module:7-27). */
static void FN(el_tables)(REAL t[5][5])
{
    for (int a = 0; a < 5; ++a)
        for (int b = 0; b < 5; ++b) t[a][b] = 0;
    t[0][0] = (REAL)1.0f;
    t[1][0] = (REAL)(9.0f / 8.0f);
    t[1][1] = (REAL)(-1.0f / 24.0f);
    t[2][0] = (REAL)(75.0f / 64.0f);
    t[2][1] = (REAL)(-25.0f / 384.0f);
    t[2][2] = (REAL)(3.0f / 640.0f);
    t[3][0] = (REAL)(1225.0f / 1024.0f);
    t[3][1] = (REAL)(-245.0f / 3072.0f);
    t[3][2] = (REAL)(49.0f / 5120.0f);
    t[3][3] = (REAL)(-5.0f / 7168.0f);
    t[4][0] = (REAL)(19845.0f / 16384.0f);
    t[4][1] = (REAL)(-735.0f / 8192.0f);
    t[4][2] = (REAL)(567.0f / 40960.0f);
    t[4][3] = (REAL)(-405.0f / 229376.0f);
    t[4][4] = (REAL)(35.0f / 294912.0f);
}

/* calux (pub:1026-1084) / caluxabs (pub:1088-1157): ux += (dt/dx) pa
 * sum a10n (u(k,i+n+s) - u(k,i-n-1+s)), s = xflag, rows k >= kstart(i)
 * (itopo in 'fsbc' mode, itopo - npzn in 'abc' mode). */
static void FN(el_calux)(int xflag, int nx, int nz, REAL dx, REAL dt, const REAL *u, REAL *ux, const REAL *pa,
                         const int *kstart, const REAL a10[5], int nth)
{
#pragma omp parallel for num_threads(nth) schedule(static) if (nth > 1)
    for (int i = 0; i < nx; ++i)
        for (int k = kstart[i]; k < nz; ++k) {
            REAL s = 0;
            for (int n = 0; n < 5; ++n) {
                const REAL d = UF(u, i + n + xflag, k) - UF(u, i - n - 1 + xflag, k);
                s = (n == 0) ? a10[0] * d : s + a10[n] * d;
            }
            UF(ux, i, k) = UF(ux, i, k) + (dt / dx) * UF(pa, i, k) * s;
        }
}

/* caluz (pub:1160-1337): zflag 0 uses u(k+n) - u(k-n-1), 10th order from
 * itopo+5, then 8th, 6th, 4th, 2nd, 2nd at itopo+4 .. itopo; zflag 1 uses
 * u(k+n+1) - u(k-n), 10th order from itopo+4, then 8th, 6th, 4th, 2nd.
 * caluzabs (pub:1341-1406, full != 0): 10th order on every row >= kstart. */
static void FN(el_caluz)(int zflag, int nx, int nz, REAL dz, REAL dt, const REAL *u, REAL *uz, const REAL *pa,
                         const int *kstart, const int *itopo, int full, REAL tab[5][5], int nth)
{
#pragma omp parallel for num_threads(nth) schedule(static) if (nth > 1)
    for (int i = 0; i < nx; ++i)
        for (int k = kstart[i]; k < nz; ++k) {
            const int d = k - itopo[i];
            int ord; /* table index: 4 = 10th ... 0 = 2nd */
            if (full) ord = 4;
            else if (zflag == 0) ord = d >= 5 ? 4 : (d == 4 ? 3 : d == 3 ? 2 : d == 2 ? 1 : 0);
            else ord = d >= 4 ? 4 : (d == 3 ? 3 : d == 2 ? 2 : d == 1 ? 1 : 0);
            REAL s = 0;
            for (int n = 0; n <= ord; ++n) {
                const REAL v = UF(u, i, k + n + zflag) - UF(u, i, k - n - 1 + zflag);
                s = (n == 0) ? tab[ord][0] * v : s + tab[ord][n] * v;
            }
            UF(uz, i, k) = UF(uz, i, k) + (dt / dz) * UF(pa, i, k) * s;
        }
}

/* calpmlbclrb (pub:895-934) / calpmlbclrtb (pub:851-892): x parts damped in
 * the left/right strips for rows k >= kstart (the right strip uses the
 * kstart of the column counted from the left edge, as written); z parts in
 * the bottom strip and, in 'abc' mode (tzn != NULL), in the npzn + 1 rows
 * itopo - npzn .. itopo of every column by tzn(k - (itopo - npzn)). */
static void FN(el_sponge)(REAL *ux, REAL *uz, const REAL *txn, const REAL *txp, const REAL *tzn, const REAL *tzp,
                          int npxn, int npxp, int npzn, int npzp, int nx, int nz, const int *kstart, const int *itopo)
{
    for (int i = 0; i <= npxn; ++i)
        for (int k = kstart[i]; k < nz; ++k) UF(ux, i, k) = UF(ux, i, k) * txn[i];
    for (int i = 0; i <= npxp; ++i)
        for (int k = kstart[i]; k < nz; ++k) UF(ux, nx - 1 - i, k) = UF(ux, nx - 1 - i, k) * txp[i];
    if (tzn)
        for (int i = 0; i < nx; ++i)
            for (int k = itopo[i] - npzn; k <= itopo[i]; ++k)
                if (k >= 0) UF(uz, i, k) = UF(uz, i, k) * tzn[k - (itopo[i] - npzn)];
    for (int i = 0; i < nx; ++i)
        for (int k = 0; k <= npzp; ++k) UF(uz, i, nz - 1 - k) = UF(uz, i, nz - 1 - k) * tzp[k];
}

/* gauss_filter_2d (pub:592-634) as one thread runs it: every cell of rows
 * 0 .. itopo(i) replaced in place, column by column, by the gc2d-weighted sum
 * of its (2 rgau - 1)^2 window clipped to the grid (not renormalised), so
 * cells already smoothed feed the later ones.  (The reference's OpenMP loop
 * over columns races on the shared neighbours; the serial order is the
 * deterministic reading.) */
static void FN(el_gauss_smooth)(int nx, int nz, int rgau, const int *itopo, const REAL *g2c, REAL *u)
{
    const int ng = 2 * rgau - 1;
    for (int i = 0; i < nx; ++i)
        for (int k = 0; k <= itopo[i]; ++k) {
            REAL sum = 0;
            int idx2 = 1;
            for (int n = i - (rgau - 1); n <= i + (rgau - 1); ++n, ++idx2) {
                int idx1 = 1;
                for (int m = k - (rgau - 1); m <= k + (rgau - 1); ++m, ++idx1)
                    if (n >= 0 && n < nx && m >= 0 && m < nz && idx2 <= ng && idx1 <= ng)
                        sum = sum + AT(u, n, m) * g2c[(size_t)(idx2 - 1) * ng + (idx1 - 1)];
            }
            AT(u, i, k) = sum;
        }
}

/* csponge (pub:796-847) for one side, outermost cell first. */
static void FN(el_csponge)(int np, REAL h, REAL Rc, REAL *spg, REAL *spgh, int shift)
{
    if (np <= 0) {
        spg[0] = 1;
        spgh[0] = 1;
        return;
    }
#if ORACLE_F32
    const float fac = acosf(Rc) / ((float)np * h);
#define COS_ cosf
#else
    const double fac = acos(Rc) / ((double)np * h);
#define COS_ cos
#endif
    for (int i = 1; i <= np + 1; ++i) {
        spg[i - 1] = COS_(fac * (REAL)(np - i + 1) * h);
        if (shift < 0) {
            spgh[i - 1] = COS_(fac * ((REAL)(np - i + 1) * h - h / 2));
            if (i == np + 1) spgh[i - 1] = 1;
        } else {
            spgh[i - 1] = COS_(fac * ((REAL)(np - i + 1) * h + h / 2));
        }
    }
#undef COS_
}

/* substagger (pub:323-547): Lame parameters, averaged buoyancies and moduli
 * at the staggered positions, then every parameter of rows above the surface
 * replaced by the surface row's (pub:507-522).  out = [5][nx*nz]: busg2e,
 * busg1e, lambmusg, lambsg, musge.  Returns -1 on a zero density. */
static int FN(el_stagger)(int nx, int nz, const float *vpi, const float *vsi, const float *rhoi, const int *itopo,
                          REAL *out)
{
    const size_t n = (size_t)nx * nz;
    REAL *w = calloc(13 * n, sizeof(REAL));
    if (!w) return -2;
    REAL *vp = w, *vs = w + n, *rho = w + 2 * n, *mu = w + 3 * n, *lamb = w + 4 * n, *musg = w + 5 * n;
    REAL *mge = w + 6 * n, *lsg = w + 7 * n, *lmsg = w + 8 * n, *b1 = w + 9 * n, *b2 = w + 10 * n;
    REAL *b1e = w + 11 * n, *b2e = w + 12 * n;
    for (size_t c = 0; c < n; ++c) {
        vp[c] = (REAL)vpi[c];
        vs[c] = (REAL)vsi[c];
        rho[c] = (REAL)rhoi[c];
    }
    for (size_t c = 0; c < n; ++c) {
        mu[c] = rho[c] * vs[c] * vs[c];
        lamb[c] = rho[c] * vp[c] * vp[c] - (REAL)2. * mu[c];
    }
    for (int i = 0; i < nx - 1; ++i)
        for (int k = 0; k < nz - 1; ++k)
            AT(b1, i, k) = (REAL)0.25 * (((AT(rho, i, k) + AT(rho, i, k + 1)) + AT(rho, i + 1, k + 1)) + AT(rho, i + 1, k));
    for (int i = 0; i < nx; ++i) AT(b1, i, nz - 1) = AT(b1, i, nz - 2);
    for (int k = 0; k < nz; ++k) AT(b1, nx - 1, k) = AT(b1, nx - 2, k);
    for (size_t c = 0; c < n; ++c) {
        if (b1[c] == 0) goto bad;
        b1[c] = (REAL)1. / b1[c];
    }
    for (int i = 0; i < nx; ++i) {
        for (int k = 1; k < nz; ++k) AT(b1e, i, k) = (REAL)0.5 * (AT(b1, i, k) + AT(b1, i, k - 1));
        AT(b1e, i, 0) = AT(b1, i, 0);
    }
    for (size_t c = 0; c < n; ++c) {
        if (rho[c] == 0) goto bad;
        b2[c] = (REAL)1. / rho[c];
    }
    for (int i = 0; i < nx - 1; ++i)
        for (int k = 0; k < nz; ++k) AT(b2e, i, k) = (REAL)0.5 * (AT(b2, i, k) + AT(b2, i + 1, k));
    for (int k = 0; k < nz; ++k) AT(b2e, nx - 1, k) = AT(b2, nx - 1, k);
    for (int i = 0; i < nx; ++i) {
        for (int k = 0; k < nz - 1; ++k) AT(musg, i, k) = (REAL)0.5 * (AT(mu, i, k + 1) + AT(mu, i, k));
        AT(musg, i, nz - 1) = AT(mu, i, nz - 1);
    }
    for (int i = 0; i < nx - 1; ++i)
        for (int k = 1; k < nz; ++k) {
            const REAL m0 = AT(musg, i, k), m1 = AT(musg, i + 1, k), m2 = AT(musg, i + 1, k - 1), m3 = AT(musg, i, k - 1);
            AT(mge, i, k) = (m0 != 0 && m3 != 0 && m1 != 0 && m2 != 0)
                                ? (REAL)4. / ((((REAL)1. / m0 + (REAL)1. / m1) + (REAL)1. / m2) + (REAL)1. / m3)
                                : 0;
        }
    for (int i = 0; i < nx; ++i) AT(mge, i, 0) = AT(musg, i, 0);
    for (int k = 0; k < nz; ++k) AT(mge, nx - 1, k) = AT(musg, nx - 1, k);
    for (int k = 0; k < nz; ++k) {
        for (int i = 0; i < nx - 1; ++i) AT(lsg, i, k) = (REAL)0.5 * (AT(lamb, i + 1, k) + AT(lamb, i, k));
        AT(lsg, nx - 1, k) = AT(lamb, nx - 1, k);
    }
    for (int i = 0; i < nx - 1; ++i)
        for (int k = 0; k < nz; ++k)
            AT(lmsg, i, k) = (REAL)0.5 * (((AT(lamb, i, k) + 2 * AT(mu, i, k)) + AT(lamb, i + 1, k)) + 2 * AT(mu, i + 1, k));
    for (int k = 0; k < nz; ++k) AT(lmsg, nx - 1, k) = AT(lamb, nx - 1, k) + 2 * AT(mu, nx - 1, k);
    /* Parameters above the topography copied from the surface row. */
    for (int i = 0; i < nx; ++i)
        for (int k = itopo[i]; k >= 1; --k) {
            REAL *arrs[13] = {vp, vs, rho, lamb, lsg, lmsg, mu, musg, mge, b1, b2, b1e, b2e};
            for (int q = 0; q < 13; ++q) AT(arrs[q], i, k - 1) = AT(arrs[q], i, k);
        }
    REAL *src[5] = {b2e, b1e, lmsg, lsg, mge};
    for (int q = 0; q < 5; ++q) memcpy(out + (size_t)q * n, src[q], n * sizeof(REAL));
    free(w);
    return 0;
bad:
    free(w);
    return -1;
}

int FN(oracle_elastic)(const oracle_elastic_args *a, REAL *rec_vx, REAL *rec_vz, REAL *snaps)
{
    const int nx = a->nx, nz = a->nz, nth = a->nthreads > 0 ? a->nthreads : 1;
    if (nx < 12 || nz < 12 || a->nsteps < 0 || a->rgau < 1) return -1;
    if (a->npxn < 0 || a->npxp < 0 || a->npzp < 0 || a->npxn + 1 > nx || a->npxp + 1 > nx || a->npzp + 1 > nz)
        return -1;
    if (a->mode != 0 && a->mode != 1) return -1;
    if (a->mode == 1 && (a->npzn < 1 || a->npzn + 1 > nz)) return -1;
    if (a->isx < 0 || a->isx >= nx || a->isz < 0 || a->isz >= nz) return -1;
    for (int i = 0; i < nx; ++i)
        if (a->itopo[i] < 0 || a->itopo[i] > nz - 6) return -1;
    for (int r = 0; r < a->nrec; ++r)
        if (a->rx[r] < 0 || a->rx[r] >= nx || a->rz[r] < 0 || a->rz[r] >= nz) return -1;

    const size_t nh = (size_t)(nx + 2 * H) * NZH, n = (size_t)nx * nz;
    int rc = -2;
    REAL *buf = calloc(19 * nh, sizeof(REAL));
    REAL *g2c = NULL, *spt = NULL;
    int *kst = NULL;
    if (!buf) return -2;
    REAL *vx = buf, *vz = buf + nh, *sxx = buf + 2 * nh, *szz = buf + 3 * nh, *sxz = buf + 4 * nh;
    REAL *vxx = buf + 5 * nh, *vxz = buf + 6 * nh, *vzx = buf + 7 * nh, *vzz = buf + 8 * nh;
    REAL *sxxx = buf + 9 * nh, *sxxz = buf + 10 * nh, *szzx = buf + 11 * nh, *szzz = buf + 12 * nh;
    REAL *sxzx = buf + 13 * nh, *sxzz = buf + 14 * nh;
    REAL *busg2e = buf + 15 * nh, *busg1e = buf + 16 * nh, *lambmusg = buf + 17 * nh, *lambsg = buf + 18 * nh;
    REAL *musge = NULL;
    const int *itopo = a->itopo;

#if ORACLE_F32
    const float dt = a->dt, dx = a->dx, dz = a->dz, f0 = a->f0;
#else
    const double dt = a->dt, dx = a->dx, dz = a->dz, f0 = a->f0;
#endif
    const REAL PI = (REAL)3.1415926f; /* the module constant pi (module:4) */

    const int abc = a->mode == 1, npzn = abc ? a->npzn : 0;
    kst = malloc((size_t)nx * sizeof(int));
    if (!kst) goto done;
    /* First updated row of every column: the surface ('fsbc'), or npzn rows
     * above it ('abc'; row -1 of the split parts is never summed). */
    for (int i = 0; i < nx; ++i) kst[i] = abc ? (itopo[i] - npzn > 0 ? itopo[i] - npzn : 0) : itopo[i];

    /* get_gauss_filter_coe(gc2d, 2.0, 2 rgau - 1, 2 rgau - 1) (pub:1470-1501). */
    const int ng = 2 * a->rgau - 1;
    g2c = malloc((size_t)ng * ng * sizeof(REAL));
    if (!g2c) goto done;
    {
        const REAL sigma = (REAL)2.0, sigma2 = sigma * sigma;
        const int cen = (ng + 1) / 2;
        REAL sum = 0;
        for (int j = 1; j <= ng; ++j)
            for (int i = 1; i <= ng; ++i) {
                const REAL dis = (REAL)(abs(j - cen) * abs(j - cen) + abs(i - cen) * abs(i - cen));
#if ORACLE_F32
                g2c[(size_t)(j - 1) * ng + (i - 1)] = 1 / (2 * PI * sigma2) * expf(-dis / (2 * sigma2));
#else
                g2c[(size_t)(j - 1) * ng + (i - 1)] = 1 / (2 * PI * sigma2) * exp(-dis / (2 * sigma2));
#endif
            }
        for (int j = 1; j <= ng; ++j)
            for (int i = 1; i <= ng; ++i) sum = sum + g2c[(size_t)(j - 1) * ng + (i - 1)];
        for (int c = 0; c < ng * ng; ++c) g2c[c] = g2c[c] / sum;
    }

    /* ---- substagger (pub:323-547), then in 'abc' mode modipapml
     * (pub:549-590): the parameters on and above the surface smoothed ---- */
    {
        REAL *st = malloc(5 * n * sizeof(REAL));
        musge = calloc(nh, sizeof(REAL));
        if (!st || !musge) {
            free(st);
            goto done;
        }
        if (FN(el_stagger)(nx, nz, a->vp, a->vs, a->rho, itopo, st) != 0) {
            free(st);
            rc = -1;
            goto done;
        }
        if (abc)
            for (int q = 0; q < 5; ++q) FN(el_gauss_smooth)(nx, nz, a->rgau, itopo, g2c, st + (size_t)q * n);
        REAL *dst[5] = {busg2e, busg1e, lambmusg, lambsg, musge};
        for (int q = 0; q < 5; ++q)
            for (int i = 0; i < nx; ++i)
                for (int k = 0; k < nz; ++k) UF(dst[q], i, k) = AT(st + (size_t)q * n, i, k);
        free(st);
    }

    REAL tab[5][5];
    FN(el_tables)(tab);
    const REAL *a10 = tab[4];

    /* csponge (pub:796-847): tables outermost first. */
    spt = malloc((size_t)2 * ((a->npxn + 1) + (a->npxp + 1) + (a->npzp + 1) + (npzn + 1)) * sizeof(REAL));
    if (!spt) goto done;
    REAL *spgxn = spt, *spg1xn = spgxn + a->npxn + 1, *spgxp = spg1xn + a->npxn + 1, *spg2xp = spgxp + a->npxp + 1;
    REAL *spgzp = spg2xp + a->npxp + 1, *spg2zp = spgzp + a->npzp + 1;
    REAL *spgzn = spg2zp + a->npzp + 1, *spg1zn = spgzn + npzn + 1;
    FN(el_csponge)(a->npxn, dx, (REAL)a->sponge_r, spgxn, spg1xn, -1);
    FN(el_csponge)(a->npxp, dx, (REAL)a->sponge_r, spgxp, spg2xp, +1);
    FN(el_csponge)(a->npzp, dz, (REAL)a->sponge_r, spgzp, spg2zp, +1);
    FN(el_csponge)(npzn, dz, (REAL)a->sponge_r, spgzn, spg1zn, -1);
    const REAL *tzn = abc ? spgzn : NULL, *tzn1 = abc ? spg1zn : NULL;
    const int isxe = a->isx + 1, isze = a->isz + 1; /* Fortran indices */

    const int nrec_t = a->nsteps - a->delay;
    int snap_next = 0;
    for (int s = 0; s < a->nsteps; ++s) {
        const int it = s + 1;
        /* addsrc (pub:1429-1466) into u with this step's wavelet. */
        REAL wv;
        if (a->wavelet) {
            wv = (REAL)a->wavelet[s];
        } else {
#if ORACLE_F32
            const float x = PI * f0 * ((float)it * dt - (float)a->delay * dt);
            wv = expf(-(x * x)) * (1 - 2 * (x * x));
#else
            const double x = PI * f0 * ((double)it * dt - (double)a->delay * dt);
            wv = exp(-(x * x)) * (1 - 2 * (x * x));
#endif
        }
#define ADDSRC(u)                                                                                          \
    do {                                                                                                   \
        int idx2 = 1;                                                                                      \
        for (int nn = isxe - (a->rgau - 1); nn <= isxe + (a->rgau - 1); ++nn, ++idx2) {                   \
            int idx1 = 1;                                                                                  \
            for (int m = isze - (a->rgau - 1); m <= isze + (a->rgau - 1); ++m, ++idx1) {                   \
                if (nn >= 1 && nn <= nx && m >= 1 && m <= nz && idx2 <= ng && idx1 <= ng)                   \
                    UF(u, nn - 1, m - 1) = wv * g2c[(size_t)(idx2 - 1) * ng + (idx1 - 1)] * dt / (dz * dz) + \
                                           UF(u, nn - 1, m - 1);                                           \
            }                                                                                              \
        }                                                                                                  \
    } while (0)

        /* ---- velocities (pri:119-144; 'abc': pri:300-318) ---- */
        FN(el_calux)(0, nx, nz, dx, dt, sxx, vxx, busg2e, kst, a10, nth);
        FN(el_caluz)(0, nx, nz, dz, dt, sxz, vxz, busg2e, kst, itopo, abc, tab, nth);
        FN(el_sponge)(vxx, vxz, spgxn, spgxp, tzn, spgzp, a->npxn, a->npxp, npzn, a->npzp, nx, nz, kst, itopo);
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k) UF(vx, i, k) = UF(vxx, i, k) + UF(vxz, i, k);
        FN(el_calux)(1, nx, nz, dx, dt, sxz, vzx, busg1e, kst, a10, nth);
        FN(el_caluz)(1, nx, nz, dz, dt, szz, vzz, busg1e, kst, itopo, abc, tab, nth);
        FN(el_sponge)(vzx, vzz, spg1xn, spg2xp, tzn1, spg2zp, a->npxn, a->npxp, npzn, a->npzp, nx, nz, kst, itopo);
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k) UF(vz, i, k) = UF(vzx, i, k) + UF(vzz, i, k);
        /* calfsbcv on both pairs of split parts (pri:147-151): vz above the
         * surface for the next step's sums. */
        if (!abc)
            for (int pass = 0; pass < 2; ++pass) {
                REAL *px = pass ? vxz : vxx, *pz = pass ? vzz : vzx;
                for (int ix = 0; ix < nx - 1; ++ix) {
                    const int izh = itopo[ix];
                    UF(pz, ix, izh - 1) = UF(pz, ix, izh) + (UF(lambsg, ix, izh) / UF(lambmusg, ix, izh)) *
                                                              (UF(px, ix + 1, izh) - UF(px, ix, izh));
                }
            }

        /* ---- tau_xx (pri:156-170; 'abc': pri:322-331, source always) ---- */
        FN(el_calux)(1, nx, nz, dx, dt, vx, sxxx, lambmusg, kst, a10, nth);
        FN(el_caluz)(0, nx, nz, dz, dt, vz, sxxz, lambsg, kst, itopo, abc, tab, nth);
        FN(el_sponge)(sxxx, sxxz, spg1xn, spg2xp, tzn, spgzp, a->npxn, a->npxp, npzn, a->npzp, nx, nz, kst, itopo);
        if (abc || isze - 1 > itopo[isxe - 1]) ADDSRC(sxxx);
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k) UF(sxx, i, k) = UF(sxxx, i, k) + UF(sxxz, i, k);

        /* ---- tau_zz (pri:172-184; 'abc': pri:333-342) ---- */
        FN(el_calux)(1, nx, nz, dx, dt, vx, szzx, lambsg, kst, a10, nth);
        FN(el_caluz)(0, nx, nz, dz, dt, vz, szzz, lambmusg, kst, itopo, abc, tab, nth);
        FN(el_sponge)(szzx, szzz, spg1xn, spg2xp, tzn, spgzp, a->npxn, a->npxp, npzn, a->npzp, nx, nz, kst, itopo);
        if (!abc)
            for (int ix = 0; ix < nx; ++ix)
                for (int iz = -1; iz <= itopo[ix]; ++iz) {
                    UF(szzz, ix, iz) = 0;
                    UF(szzx, ix, iz) = 0;
                }
        ADDSRC(szzx);
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k) UF(szz, i, k) = UF(szzx, i, k) + UF(szzz, i, k);

        /* ---- tau_xz (pri:186-196; 'abc': pri:344-352) ---- */
        FN(el_calux)(0, nx, nz, dx, dt, vz, sxzx, musge, kst, a10, nth);
        FN(el_caluz)(1, nx, nz, dz, dt, vx, sxzz, musge, kst, itopo, abc, tab, nth);
        FN(el_sponge)(sxzx, sxzz, spgxn, spgxp, tzn1, spg2zp, a->npxn, a->npxp, npzn, a->npzp, nx, nz, kst, itopo);
        if (!abc)
            for (int ix = 0; ix < nx; ++ix) {
                const int izh = itopo[ix];
                UF(sxzx, ix, izh - 1) = -UF(sxzx, ix, izh);
                UF(sxzz, ix, izh - 1) = -UF(sxzz, ix, izh);
            }
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k) UF(sxz, i, k) = UF(sxzx, i, k) + UF(sxzz, i, k);
#undef ADDSRC

        /* ---- records (pri:199-210) and snapshots ---- */
        const int col = s - a->delay;
        if (col >= 0 && col < nrec_t)
            for (int r = 0; r < a->nrec; ++r) {
                if (rec_vx) rec_vx[(size_t)r * nrec_t + col] = UF(vx, a->rx[r], a->rz[r]);
                if (rec_vz) rec_vz[(size_t)r * nrec_t + col] = UF(vz, a->rx[r], a->rz[r]);
            }
        while (snaps && snap_next < a->nsnap && a->snap_steps[snap_next] == s) {
            REAL *dst = snaps + (size_t)snap_next * 5 * n;
            for (int i = 0; i < nx; ++i)
                for (int k = 0; k < nz; ++k) {
                    AT(dst, i, k) = UF(vx, i, k);
                    AT(dst + n, i, k) = UF(vz, i, k);
                    AT(dst + 2 * n, i, k) = UF(sxx, i, k);
                    AT(dst + 3 * n, i, k) = UF(szz, i, k);
                    AT(dst + 4 * n, i, k) = UF(sxz, i, k);
                }
            ++snap_next;
        }
    }
    rc = 0;
done:
    free(buf);
    free(musge);
    free(g2c);
    free(spt);
    free(kst);
    return rc;
}

#undef H
#undef NZH
#undef UF
#undef AT
