/*
 * Body of the visco-elastic rotated-staggered-grid oracle, included twice by
 * visco.c:
 *   REAL = float,  ORACLE_F32 = 1 : single precision, expressions evaluated
 *                                   in the order the Fortran source writes them
 *   REAL = double, ORACLE_F32 = 0 : the same steps in double
 * FN(x) names the instantiated function.  TEST INFRASTRUCTURE ONLY.
 *
 * Restates forward/Q_forward/Q_multishot_fd_subroutine_fsbc_zy_rotated.f90:
 *   substagger        :434-469   density average and buoyancy
 *   pcalpas           :473-551   moduli and the tau-method (L = 1 as written;
 *                                L > 1 by Blanch's tau-method, see visco.c)
 *   csponge           :555-605   cosine sponge tables
 *   calpmlbclrtb      :609-654   split-field sponge on all four sides
 *   extrapolation_one_step_10nd_absbc :695-839, and its callees
 *   calderxabs/calderzabs :843-966, calpvabs :969-990, calpstr* :993-1132,
 *   calpr* :1135-1269, calsum* :1273-1337, addsrc :1340-1392,
 *   get_gauss_filter_coe :1398-1444
 * and the time loop with recording of
 * forward/Q_forward/Q_multishot_fd_main_fsbc_zy_rotated.f90:274-280,436-459.
 *
 * Layout: every field is [nx][nz] (x rows, z fastest; the Fortran a(k,i)
 * with k = z).  Full fields carry H = 5 zero cells on each side, as the
 * Fortran's -4:nz+5 bounds do; split fields and derivatives do not.
 * Fortran index (k, i) = C index (i - 1, k - 1).
 */

#define H 5
#define NZH (nz + 2 * H)
#define UF(u, i, k) (u)[(size_t)((i) + H) * NZH + (size_t)((k) + H)]
#define AT(u, i, k) (u)[(size_t)(i) * nz + (size_t)(k)]

/* calderxabs / calderzabs: dir 0 = x, 1 = z; shift 0 = xflag/zflag 0,
 * shift 1 = xflag/zflag 1. */
static void FN(deriv)(int dir, int shift, const REAL *u, REAL *der, REAL h, const REAL *a, int R,
                      int nx, int nz, int nthreads)
{
    const REAL fac = (REAL)0.5 / h;
#pragma omp parallel for num_threads(nthreads) schedule(static) if (nthreads > 1)
    for (int i = 0; i < nx; ++i) {
        for (int k = 0; k < nz; ++k) {
            REAL s = 0;
            for (int n = 0; n < R; ++n) {
                REAL t;
                if (shift == 0) {
                    const REAL u1 = UF(u, i + n, k + n), u2 = UF(u, i - n - 1, k - n - 1);
                    const REAL u3 = UF(u, i + n, k - n - 1), u4 = UF(u, i - n - 1, k + n);
                    t = dir == 0 ? ((u1 - u2) + u3) - u4 : ((u1 - u2) - u3) + u4;
                } else {
                    const REAL u1 = UF(u, i + 1 + n, k + 1 + n), u2 = UF(u, i - n, k - n);
                    const REAL u3 = UF(u, i + 1 + n, k - n), u4 = UF(u, i - n, k + 1 + n);
                    t = dir == 0 ? ((u1 - u2) + u3) - u4 : ((u1 - u2) - u3) + u4;
                }
                s = (n == 0) ? a[0] * t : s + a[n] * t;
            }
            AT(der, i, k) = fac * s;
        }
    }
}

/* calpvabs: u = u + dt*pa*der. */
static void FN(pv)(REAL *u, const REAL *pa, const REAL *der, REAL dt, size_t n, int nthreads)
{
#pragma omp parallel for num_threads(nthreads) schedule(static) if (nthreads > 1)
    for (size_t c = 0; c < n; ++c) u[c] = u[c] + dt * pa[c] * der[c];
}

/* calpmlbclrtb: x parts damped in the left/right strips (all depths), z parts
 * in the top/bottom strips (all columns).  tx_n[i], i = 0..npxn, is the table
 * value of Fortran index i + 1; likewise for the other three. */
static void FN(sponge)(REAL *ux, REAL *uz, const REAL *txn, const REAL *txp, const REAL *tzn,
                       const REAL *tzp, int npxn, int npxp, int npzn, int npzp, int nx, int nz)
{
    for (int i = 0; i <= npxn; ++i)
        for (int k = 0; k < nz; ++k) AT(ux, i, k) = AT(ux, i, k) * txn[i];
    for (int i = 0; i <= npxp; ++i)
        for (int k = 0; k < nz; ++k) AT(ux, nx - 1 - i, k) = AT(ux, nx - 1 - i, k) * txp[i];
    for (int i = 0; i < nx; ++i)
        for (int k = 0; k <= npzn; ++k) AT(uz, i, k) = AT(uz, i, k) * tzn[k];
    for (int i = 0; i < nx; ++i)
        for (int k = 0; k <= npzp; ++k) AT(uz, i, nz - 1 - k) = AT(uz, i, nz - 1 - k) * tzp[k];
}

/* csponge for one side: spg(i) = cos(fac*float(np-i+1)*h) and the half-cell
 * variant (shift = -1: spg1 with -h/2 and spg1(np+1) = 1; shift = +1: spg2
 * with +h/2). */
static void FN(csponge_side)(int np, REAL h, REAL Rc, REAL *spg, REAL *spgh, int shift)
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

/* get_gauss_filter_coe (Q sub:1398-1444): g2c(i,j), i = 1..n1, j = 1..n2,
 * column-major, 1/(2 pi sigma^2) exp(-dis / (2 sigma^2)) with dis the squared
 * index distance to ((n1+1)/2, (n2+1)/2) (integer division), normalised by
 * its sum accumulated j-outer, i-inner. */
static void FN(gauss_filter)(int n1, int n2, REAL sigma, REAL *g2c)
{
#if ORACLE_F32
    const float PI = 3.1415926f;
#else
    const double PI = 3.1415926;
#endif
    const REAL sigma2 = sigma * sigma;
    const int ixcen = (n1 + 1) / 2, izcen = (n2 + 1) / 2;
    REAL sum = 0;
    for (int j = 1; j <= n2; ++j)
        for (int i = 1; i <= n1; ++i) {
            const REAL dis = (REAL)(abs(j - izcen) * abs(j - izcen) + abs(i - ixcen) * abs(i - ixcen));
#if ORACLE_F32
            g2c[(size_t)(j - 1) * n1 + (i - 1)] = 1 / (2 * PI * sigma2) * expf(-dis / (2 * sigma2));
#else
            g2c[(size_t)(j - 1) * n1 + (i - 1)] = 1 / (2 * PI * sigma2) * exp(-dis / (2 * sigma2));
#endif
        }
    for (int j = 1; j <= n2; ++j)
        for (int i = 1; i <= n1; ++i) sum = sum + g2c[(size_t)(j - 1) * n1 + (i - 1)];
    for (size_t c = 0; c < (size_t)n1 * n2; ++c) g2c[c] = g2c[c] / sum;
}

/* pcalpas' relaxation times (:506-542).  One mechanism (L = 1), as written:
 * tau_sigma = 1 / ((w_up - w_low) / 2) and Blanch's closed form
 *   tau = (1/Q') (ln(1 + w^2 ts^2)|) / ((atan(w ts) - w ts / (1 + w^2 ts^2))|),
 * evaluated between w_low and w_up, Q' = Q * q_scale (the reference uses
 * 18/20).  L > 1: tau_sigma_l = 1 / w_l with w_l log-spaced over the band and
 * the least-squares tau of Blanch, Robertsson & Symes (1995):
 *   tau = (1/Q') sum_l I0_l / (sum_l I1_l + 2 sum_{k<l} I2_kl).
 * Band: w_low = 2 pi f_low (reference f_low = 2 Hz), w_up = 2 pi f_high
 * (reference 2.5 f0); f_low, f_high <= 0 select the reference values.
 * tau may be NULL (relaxation times only). */
void FN(oracle_visco_tau)(int L, double f0_, double q_, double q_scale_, double f_low_, double f_high_,
                          REAL *tsig, REAL *tau)
{
#if ORACLE_F32
    const float PI = 3.1415926f;
#define LOG_ logf
#define ATAN_ atanf
#define POW_ powf
#else
    const double PI = 3.1415926;
#define LOG_ log
#define ATAN_ atan
#define POW_ pow
#endif
    const REAL f0 = (REAL)f0_;
    const REAL flow = (f_low_ > 0 ? (REAL)f_low_ : (REAL)2) * 2 * PI;
    const REAL fup = (f_high_ > 0 ? (REAL)f_high_ : (REAL)2.5 * f0) * 2 * PI;
    const REAL q = (REAL)q_ * (REAL)q_scale_;
    if (L == 1) {
        const REAL tss = (REAL)1.0 / ((fup - flow) / (REAL)2.0);
        tsig[0] = tss;
        if (tau) {
            const REAL nlow = LOG_(tss * tss * (flow * flow) + 1);
            const REAL nup = LOG_(tss * tss * (fup * fup) + 1);
            const REAL dlow = ATAN_(flow * tss) - (flow * tss) / (1 + tss * tss * (flow * flow));
            const REAL dup = ATAN_(fup * tss) - (fup * tss) / (1 + tss * tss * (fup * fup));
            *tau = (dup - dlow != 0) ? (REAL)1.0 / q * (nup - nlow) / (dup - dlow) : 0;
        }
        return;
    }
    for (int l = 0; l < L; ++l) tsig[l] = 1 / (flow * POW_(fup / flow, (REAL)l / (REAL)(L - 1)));
    if (!tau) return;
    REAL s0 = 0, s1 = 0, s2 = 0;
    for (int l = 0; l < L; ++l) {
        const REAL t = tsig[l];
        s0 += (LOG_(1 + fup * fup * t * t) - LOG_(1 + flow * flow * t * t)) / (2 * t);
        s1 += ((ATAN_(fup * t) - fup * t / (1 + fup * fup * t * t)) -
               (ATAN_(flow * t) - flow * t / (1 + flow * flow * t * t))) / (2 * t);
        for (int k = 0; k < l; ++k) {
            const REAL tk = tsig[k];
            s2 += tk * t / (t * t - tk * tk) *
                  ((ATAN_(fup * tk) / tk - ATAN_(fup * t) / t) - (ATAN_(flow * tk) / tk - ATAN_(flow * t) / t));
        }
    }
    *tau = s0 / (s1 + 2 * s2) / q;
#undef LOG_
#undef ATAN_
#undef POW_
}

int FN(oracle_visco)(const oracle_visco_args *a, REAL *rec_zz, REAL *rec_xx, REAL *snaps)
{
    const int nx = a->nx, nz = a->nz, R = a->radius, L = a->nmech;
    const int nth = a->nthreads > 0 ? a->nthreads : 1;
    if (nx < 2 || nz < 2 || R < 1 || R > 5 || L < 1 || L > ORACLE_MAX_MECH || a->nsteps < 0) return -1;
    if (a->npxn < 0 || a->npxp < 0 || a->npzn < 0 || a->npzp < 0) return -1;
    if (a->npxn + 1 > nx || a->npxp + 1 > nx || a->npzn + 1 > nz || a->npzp + 1 > nz) return -1;
    if (a->isx < 0 || a->isx >= nx || a->isz < 0 || a->isz >= nz) return -1;
    for (int r = 0; r < a->nrec; ++r)
        if (a->rx[r] < 0 || a->rx[r] >= nx || a->rz[r] < 0 || a->rz[r] >= nz) return -1;
    if (!a->vp || !a->rho || !a->qp) return -1;

    const size_t n = (size_t)nx * nz, nh = (size_t)(nx + 2 * H) * (nz + 2 * H);
    int rc = -2;
    /* Full fields with halos. */
    REAL *vx = calloc(nh, sizeof(REAL)), *vz = calloc(nh, sizeof(REAL));
    REAL *sxx = calloc(nh, sizeof(REAL)), *szz = calloc(nh, sizeof(REAL)), *sxz = calloc(nh, sizeof(REAL));
    /* Split parts: velocity (x, z), stress (x, z, r), memory (x, z, r) per mechanism. */
    REAL *vxx = calloc(n, sizeof(REAL)), *vxz = calloc(n, sizeof(REAL));
    REAL *vzx = calloc(n, sizeof(REAL)), *vzz = calloc(n, sizeof(REAL));
    REAL *sp = calloc(9 * n, sizeof(REAL));              /* sxx{x,z,r}, szz{..}, sxz{..} */
    REAL *rs = calloc((size_t)L * 12 * n, sizeof(REAL));  /* per mech: rxx, rzz, rxz sums, then 9 parts */
    REAL *der = calloc(4 * n, sizeof(REAL));
    /* Model. */
    REAL *rhoe = malloc(n * sizeof(REAL)), *busg = malloc(n * sizeof(REAL));
    REAL *mu = malloc(n * sizeof(REAL)), *lambmu = malloc(n * sizeof(REAL));
    REAL *tpss = malloc((size_t)L * n * sizeof(REAL)), *tsss = malloc((size_t)L * n * sizeof(REAL));
    REAL *tpsum = malloc(n * sizeof(REAL)), *tssum = malloc(n * sizeof(REAL));
    REAL *g2c = NULL;
    if (!vx || !vz || !sxx || !szz || !sxz || !vxx || !vxz || !vzx || !vzz || !sp || !rs || !der || !rhoe ||
        !busg || !mu || !lambmu || !tpss || !tsss || !tpsum || !tssum)
        goto done;

#if ORACLE_F32
    const float dt = a->dt, dx = a->dx, dz = a->dz, f0 = a->f0;
    const float PI = 3.1415926f; /* module constant, Q sub:4 */
#else
    const double dt = a->dt, dx = a->dx, dz = a->dz, f0 = a->f0;
    const double PI = 3.1415926;
#endif
    REAL coef[5];
    for (int m = 0; m < R; ++m) coef[m] = (REAL)a->coef[m];

    /* substagger (:434-469): 4-point average in place (the loop reads only
     * not-yet-updated neighbours), last row and column replicated, 1/rho. */
    for (size_t c = 0; c < n; ++c) rhoe[c] = (REAL)a->rho[c];
    for (int i = 0; i < nx - 1; ++i)
        for (int k = 0; k < nz - 1; ++k)
            AT(rhoe, i, k) = (REAL)0.25 * (((AT(rhoe, i, k) + AT(rhoe, i, k + 1)) + AT(rhoe, i + 1, k + 1)) +
                                           AT(rhoe, i + 1, k));
    for (int i = 0; i < nx; ++i) AT(rhoe, i, nz - 1) = AT(rhoe, i, nz - 2);
    for (int k = 0; k < nz; ++k) AT(rhoe, nx - 1, k) = AT(rhoe, nx - 2, k);
    for (size_t c = 0; c < n; ++c) {
        if (rhoe[c] == 0) goto done_bad;
        busg[c] = (REAL)1 / rhoe[c];
    }

    /* pcalpas (:473-551): moduli from the averaged density; tau per cell. */
    {
        REAL tsig[ORACLE_MAX_MECH];
        FN(oracle_visco_tau)(L, f0, 1.0f, a->q_scale, a->f_low, a->f_high, tsig, NULL);
        for (size_t c = 0; c < n; ++c) {
            const REAL vpc = (REAL)a->vp[c], vsc = a->vs ? (REAL)a->vs[c] : 0;
            mu[c] = rhoe[c] * (vsc * vsc);
            lambmu[c] = rhoe[c] * (vpc * vpc);
            REAL taup, taus;
            FN(oracle_visco_tau)(L, f0, a->qp[c], a->q_scale, a->f_low, a->f_high, tsig, &taup);
            FN(oracle_visco_tau)(L, f0, a->qs ? a->qs[c] : a->qp[c], a->q_scale, a->f_low, a->f_high, tsig, &taus);
            for (int l = 0; l < L; ++l) {
                const REAL tep = taup * tsig[l] + tsig[l], tes = taus * tsig[l] + tsig[l];
                tpss[(size_t)l * n + c] = tep / tsig[l];
                tsss[(size_t)l * n + c] = tes / tsig[l];
            }
            /* Unrelaxed factor: taupstinss for one mechanism (:1020), and
             * 1 + sum_l (tau_eps_l / tau_sigma_l - 1) in general. */
            REAL sp_ = tpss[c], ss_ = tsss[c];
            for (int l = 1; l < L; ++l) {
                sp_ = sp_ + (tpss[(size_t)l * n + c] - 1);
                ss_ = ss_ + (tsss[(size_t)l * n + c] - 1);
            }
            tpsum[c] = sp_;
            tssum[c] = ss_;
        }
    }
    REAL rtau[ORACLE_MAX_MECH];
    {
        REAL tsig[ORACLE_MAX_MECH];
        FN(oracle_visco_tau)(L, f0, 1.0f, a->q_scale, a->f_low, a->f_high, tsig, NULL);
        for (int l = 0; l < L; ++l) rtau[l] = (REAL)1 / tsig[l];
    }

    /* csponge (:555-605). */
    REAL spgxn[512], spgxp[512], spgzn[512], spgzp[512], spg1xn[512], spg2xp[512], spg1zn[512], spg2zp[512];
    if (a->npxn >= 511 || a->npxp >= 511 || a->npzn >= 511 || a->npzp >= 511) goto done_bad;
    FN(csponge_side)(a->npxn, dx, (REAL)a->sponge_r, spgxn, spg1xn, -1);
    FN(csponge_side)(a->npzn, dz, (REAL)a->sponge_r, spgzn, spg1zn, -1);
    FN(csponge_side)(a->npxp, dx, (REAL)a->sponge_r, spgxp, spg2xp, +1);
    FN(csponge_side)(a->npzp, dz, (REAL)a->sponge_r, spgzp, spg2zp, +1);
    /* The Fortran tables run from the outer edge inwards (index 1 = outermost
     * cell): left strip row i uses spgxn(i+1); right strip row nx-1-i uses
     * spgxp(i+1); likewise in z. */

    /* get_gauss_filter_coe (:1398-1444) over the whole grid: n1 = nz, n2 = nx. */
    const int n1 = nz, n2 = nx;
    g2c = malloc(n * sizeof(REAL));
    if (!g2c) goto done;
    FN(gauss_filter)(n1, n2, (REAL)a->src_sigma, g2c);
    /* addsrc's placement (:1360-1389): hw = nint((n+1)/2.0); the grid column
     * n2 index idx2 = n - (isxe - (hw2 - 1)) + 1, idx1 likewise in z. */
    const int hw1 = (int)lround((n1 + 1) / 2.0), hw2 = (int)lround((n2 + 1) / 2.0);
    const int isxe = a->isx + 1, isze = a->isz + 1; /* Fortran indices */

    /* Time loop (main:436-459). */
    const int nrec_t = a->nsteps - a->rec_delay;
    int snap_next = 0;
    REAL *sxxx = sp, *sxxz = sp + n, *sxxr = sp + 2 * n;
    REAL *szzx = sp + 3 * n, *szzz = sp + 4 * n, *szzr = sp + 5 * n;
    REAL *sxzx = sp + 6 * n, *sxzz = sp + 7 * n, *sxzr = sp + 8 * n;
    REAL *d1 = der, *d2 = der + n, *d3 = der + 2 * n, *d4 = der + 3 * n;
    if (a->init) {
        /* Initial state in the x parts of the split fields (the z and r parts
         * zero), mirrored into the full fields. */
        const float *I = a->init;
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k) {
                const size_t c = (size_t)i * nz + k;
                sxxx[c] = UF(sxx, i, k) = (REAL)I[c];
                szzx[c] = UF(szz, i, k) = (REAL)I[n + c];
                sxzx[c] = UF(sxz, i, k) = (REAL)I[2 * n + c];
                vxx[c] = UF(vx, i, k) = (REAL)I[3 * n + c];
                vzx[c] = UF(vz, i, k) = (REAL)I[4 * n + c];
            }
    }
    for (int s = 0; s < a->nsteps; ++s) {
        const int it = a->it0 + s;
        /* ---- velocities from the old stresses (:747-773) ---- */
        FN(deriv)(0, 0, sxx, d1, dx, coef, R, nx, nz, nth);
        FN(pv)(vxx, busg, d1, dt, n, nth);
        FN(deriv)(1, 0, sxz, d2, dz, coef, R, nx, nz, nth);
        FN(pv)(vxz, busg, d2, dt, n, nth);
        FN(sponge)(vxx, vxz, spgxn, spgxp, spgzn, spgzp, a->npxn, a->npxp, a->npzn, a->npzp, nx, nz);
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k) UF(vx, i, k) = AT(vxx, i, k) + AT(vxz, i, k);
        FN(deriv)(0, 0, sxz, d1, dx, coef, R, nx, nz, nth);
        FN(pv)(vzx, busg, d1, dt, n, nth);
        FN(deriv)(1, 0, szz, d2, dz, coef, R, nx, nz, nth);
        FN(pv)(vzz, busg, d2, dt, n, nth);
        FN(sponge)(vzx, vzz, spgxn, spgxp, spgzn, spgzp, a->npxn, a->npxp, a->npzn, a->npzp, nx, nz);
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k) UF(vz, i, k) = AT(vzx, i, k) + AT(vzz, i, k);

        /* ---- derivatives of the new velocities (:779-782) ---- */
        FN(deriv)(0, 1, vx, d1, dx, coef, R, nx, nz, nth);
        FN(deriv)(0, 1, vz, d2, dx, coef, R, nx, nz, nth);
        FN(deriv)(1, 1, vx, d3, dz, coef, R, nx, nz, nth);
        FN(deriv)(1, 1, vz, d4, dz, coef, R, nx, nz, nth);

        /* Old memory variables summed over mechanisms (the stresses take them
         * before calpr* updates them). */
        /* ---- stresses (:786-810), calpstrxx :1015-1030, calpstrzz, calpstrxz ---- */
#pragma omp parallel for num_threads(nth) schedule(static) if (nth > 1)
        for (size_t c = 0; c < n; ++c) {
            REAL rxx_ = rs[c], rzz_ = rs[n + c], rxz_ = rs[2 * n + c];
            for (int l = 1; l < L; ++l) {
                const REAL *rl = rs + (size_t)l * 12 * n;
                rxx_ = rxx_ + rl[c];
                rzz_ = rzz_ + rl[n + c];
                rxz_ = rxz_ + rl[2 * n + c];
            }
            const REAL lt = lambmu[c] * tpsum[c], mt2 = (REAL)2.0 * mu[c] * tssum[c], mt = mu[c] * tssum[c];
            sxxx[c] = sxxx[c] + dt * (lt * d1[c]);
            sxxz[c] = sxxz[c] + dt * ((lt - mt2) * d4[c]);
            sxxr[c] = sxxr[c] + dt * rxx_;
            szzx[c] = szzx[c] + dt * ((lt - mt2) * d1[c]);
            szzz[c] = szzz[c] + dt * (lt * d4[c]);
            szzr[c] = szzr[c] + dt * rzz_;
            sxzx[c] = sxzx[c] + dt * (mt * d2[c]);
            sxzz[c] = sxzz[c] + dt * (mt * d3[c]);
            sxzr[c] = sxzr[c] + dt * rxz_;
        }
        FN(sponge)(sxxx, sxxz, spg1xn, spg2xp, spg1zn, spg2zp, a->npxn, a->npxp, a->npzn, a->npzp, nx, nz);
        FN(sponge)(szzx, szzz, spg1xn, spg2xp, spg1zn, spg2zp, a->npxn, a->npxp, a->npzn, a->npzp, nx, nz);
        FN(sponge)(sxzx, sxzz, spg1xn, spg2xp, spg1zn, spg2zp, a->npxn, a->npxp, a->npzn, a->npzp, nx, nz);
        /* addsrc (:1340-1392) into the x parts of sxx and szz. */
        {
            REAL wv;
            if (a->wavelet) {
                wv = (REAL)a->wavelet[s];
            } else {
#if ORACLE_F32
                const float x = PI * f0 * ((float)it * dt - 1.2f / f0);
                wv = expf(-(x * x)) * (1 - 2 * (x * x));
#else
                const double x = PI * f0 * ((double)it * dt - 1.2 / f0);
                wv = exp(-(x * x)) * (1 - 2 * (x * x));
#endif
            }
            int idx2 = 1;
            for (int nn = isxe - (hw2 - 1); nn <= isxe + (hw2 - 1); ++nn, ++idx2) {
                if (nn < 1 || nn > nx || idx2 > n2) continue;
                int idx1 = 1;
                for (int m = isze - (hw1 - 1); m <= isze + (hw1 - 1); ++m, ++idx1) {
                    if (m < 1 || m > nz || idx1 > n1) continue;
                    const REAL g = g2c[(size_t)(idx2 - 1) * n1 + (idx1 - 1)];
                    if (g == 0) continue; /* adds +0: no change */
                    const REAL add = wv * g * dt / (dz * dz);
                    AT(sxxx, nn - 1, m - 1) = add + AT(sxxx, nn - 1, m - 1);
                    AT(szzx, nn - 1, m - 1) = add + AT(szzx, nn - 1, m - 1);
                }
            }
        }
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k) {
                UF(sxx, i, k) = (AT(sxxx, i, k) + AT(sxxz, i, k)) + AT(sxxr, i, k);
                UF(szz, i, k) = (AT(szzx, i, k) + AT(szzz, i, k)) + AT(szzr, i, k);
                UF(sxz, i, k) = (AT(sxzx, i, k) + AT(sxzz, i, k)) + AT(sxzr, i, k);
            }

        /* ---- memory variables (:814-836), calprxx :1157-1172, calprzz, calprxz ---- */
        for (int l = 0; l < L; ++l) {
            REAL *rl = rs + (size_t)l * 12 * n;
            REAL *rxx = rl, *rzz = rl + n, *rxz = rl + 2 * n;
            REAL *rxxx = rl + 3 * n, *rxxz = rl + 4 * n, *rxxr = rl + 5 * n;
            REAL *rzzx = rl + 6 * n, *rzzz = rl + 7 * n, *rzzr = rl + 8 * n;
            REAL *rxzx = rl + 9 * n, *rxzz = rl + 10 * n, *rxzr = rl + 11 * n;
            const REAL *tp = tpss + (size_t)l * n, *ts = tsss + (size_t)l * n;
            const REAL cr = dt * (-(REAL)1.0 * rtau[l]);
#pragma omp parallel for num_threads(nth) schedule(static) if (nth > 1)
            for (size_t c = 0; c < n; ++c) {
                const REAL lp = lambmu[c] * (tp[c] - 1), ms = (REAL)2.0 * mu[c] * (ts[c] - 1);
                const REAL m1 = mu[c] * (ts[c] - 1);
                rxxx[c] = rxxx[c] + cr * (lp * d1[c]);
                rxxz[c] = rxxz[c] + cr * ((lp - ms) * d4[c]);
                rxxr[c] = rxxr[c] + cr * rxx[c];
                rzzz[c] = rzzz[c] + cr * (lp * d4[c]);
                rzzx[c] = rzzx[c] + cr * ((lp - ms) * d1[c]);
                rzzr[c] = rzzr[c] + cr * rzz[c];
                rxzx[c] = rxzx[c] + cr * (m1 * d2[c]);
                rxzz[c] = rxzz[c] + cr * (m1 * d3[c]);
                rxzr[c] = rxzr[c] + cr * rxz[c];
            }
            FN(sponge)(rxxx, rxxz, spg1xn, spg2xp, spg1zn, spg2zp, a->npxn, a->npxp, a->npzn, a->npzp, nx, nz);
            FN(sponge)(rzzx, rzzz, spg1xn, spg2xp, spg1zn, spg2zp, a->npxn, a->npxp, a->npzn, a->npzp, nx, nz);
            FN(sponge)(rxzx, rxzz, spg1xn, spg2xp, spg1zn, spg2zp, a->npxn, a->npxp, a->npzn, a->npzp, nx, nz);
            for (size_t c = 0; c < n; ++c) {
                rxx[c] = (rxxx[c] + rxxz[c]) + rxxr[c];
                rzz[c] = (rzzx[c] + rzzz[c]) + rzzr[c];
                rxz[c] = (rxzx[c] + rxzz[c]) + rxzr[c];
            }
        }

        /* ---- recording (main:450-459) and snapshots ---- */
        const int col = s - a->rec_delay;
        if (col >= 0 && col < nrec_t) {
            for (int r = 0; r < a->nrec; ++r) {
                if (rec_zz) rec_zz[(size_t)r * nrec_t + col] = UF(szz, a->rx[r], a->rz[r]);
                if (rec_xx) rec_xx[(size_t)r * nrec_t + col] = UF(sxx, a->rx[r], a->rz[r]);
            }
        }
        while (snaps && snap_next < a->nsnap && a->snap_steps[snap_next] == s) {
            REAL *dst = snaps + (size_t)snap_next * 5 * n;
            for (int i = 0; i < nx; ++i)
                for (int k = 0; k < nz; ++k) {
                    AT(dst, i, k) = UF(sxx, i, k);
                    AT(dst + n, i, k) = UF(szz, i, k);
                    AT(dst + 2 * n, i, k) = UF(sxz, i, k);
                    AT(dst + 3 * n, i, k) = UF(vx, i, k);
                    AT(dst + 4 * n, i, k) = UF(vz, i, k);
                }
            ++snap_next;
        }
    }
    rc = 0;
    goto done;
done_bad:
    rc = -1;
done:
    free(vx); free(vz); free(sxx); free(szz); free(sxz);
    free(vxx); free(vxz); free(vzx); free(vzz);
    free(sp); free(rs); free(der);
    free(rhoe); free(busg); free(mu); free(lambmu); free(tpss); free(tsss); free(tpsum); free(tssum);
    free(g2c);
    return rc;
}

#undef H
#undef NZH
#undef UF
#undef AT
