/*
 * Body of the acoustic oracle, included twice by acoustic.c:
 *   REAL = float,  ORACLE_F32 = 1 : the reference's own arithmetic
 *   REAL = double, ORACLE_F32 = 0 : the same steps in double
 * FN(x) names the instantiated function.
 *
 * Restates forward/Acoustic_FD_by_ss/src/Acoustic_1order_FD.cpp:237-380
 * (the time loop) step by step.  TEST INFRASTRUCTURE ONLY.
 */

/* Reference expression `Vx2-(1.0/DEN)*(dt/dx)*(S)` (:259, :276): the quotient
 * 1.0/DEN is double, dt/dx is a float quotient, S is a float sum. */
#if ORACLE_F32
#define VEL_UPDATE(v2, den, dtd, s) \
    ((float)((double)(v2) - (1.0 / (double)(den)) * (double)(dtd) * (double)(s)))
/* Reference `P2-DEN*Vp*Vp*dt*((1.0/dx)*(Sx)+(1.0/dz)*(Sz))` (:293-311):
 * DEN*Vp*Vp*dt is a float product, the bracket is double. */
#define P_UPDATE(p2, den, vp, dt, dx, dz, sx, sz)                               \
    ((float)((double)(p2) - (double)((den) * (vp) * (vp) * (dt)) *             \
             ((1.0 / (double)(dx)) * (double)(sx) + (1.0 / (double)(dz)) * (double)(sz))))
#else
#define VEL_UPDATE(v2, den, dtd, s) ((v2) - (1.0 / (double)(den)) * (dtd) * (s))
#define P_UPDATE(p2, den, vp, dt, dx, dz, sx, sz) \
    ((p2) - (double)(den) * (vp) * (vp) * (dt) * ((1.0 / (dx)) * (sx) + (1.0 / (dz)) * (sz)))
#endif

int FN(oracle_acoustic)(const oracle_acoustic_args *a, REAL *record, REAL *snaps,
                        REAL *state, REAL *pml)
{
    const int nx = a->nx, nz = a->nz, R = a->radius;
    if (nx <= 2 * R || nz <= 2 * R || R < 1 || R > 8 || a->nt < 0 || a->k0 < 0)
        return -1;
    if (a->sx < 0 || a->sx >= nx || a->sz < 0 || a->sz >= nz) return -1;
    for (int r = 0; r < a->nrec; ++r)
        if (a->rx[r] < 0 || a->rx[r] >= nx || a->rz[r] < 0 || a->rz[r] >= nz) return -1;

    const size_t n = (size_t)nx * nz;
    REAL *p1 = (REAL *)calloc(n, sizeof(REAL)), *p2 = (REAL *)calloc(n, sizeof(REAL));
    REAL *vx1 = (REAL *)calloc(n, sizeof(REAL)), *vx2 = (REAL *)calloc(n, sizeof(REAL));
    REAL *vz1 = (REAL *)calloc(n, sizeof(REAL)), *vz2 = (REAL *)calloc(n, sizeof(REAL));
    float *absorb = NULL;
    REAL *cpml = NULL; /* [4][n] psi fields + [4][max(nx,nz)] profiles */
    if (!p1 || !p2 || !vx1 || !vx2 || !vz1 || !vz2) goto fail;

    /* Coefficients as the reference holds them: float constants (:23,38-45). */
    REAL C[8];
    for (int m = 0; m < R; ++m) C[m] = (REAL)a->coef[m];

#if ORACLE_F32
    const float dt = a->dt, dx = a->dx, dz = a->dz;
    const float dtdx = dt / dx, dtdz = dt / dz; /* float quotients, :259,:276 */
#else
    const double dt = a->dt, dx = a->dx, dz = a->dz;
    const double dtdx = dt / dx, dtdz = dt / dz;
#endif

    if (a->boundary == 0) {
        if (a->absorb) {
            absorb = (float *)a->absorb;
        } else {
            absorb = (float *)malloc(n * sizeof(float));
            if (!absorb) goto fail;
            oracle_sponge_table(nx, nz, a->sponge, absorb);
        }
    } else {
        /* C-PML (no reference counterpart): memory fields psi_px, psi_pz at the
         * velocity nodes, psi_vx, psi_vz at the pressure nodes, and the
         * recursive-convolution profiles b, a on integer and half nodes. */
        const int mx = nx > nz ? nx : nz;
        cpml = (REAL *)calloc(4 * n + 8 * (size_t)mx, sizeof(REAL));
        if (!cpml) goto fail;
    }
    REAL *psi_px = cpml, *psi_pz = cpml ? cpml + n : NULL;
    REAL *psi_vx = cpml ? cpml + 2 * n : NULL, *psi_vz = cpml ? cpml + 3 * n : NULL;
    REAL *bx_i = NULL, *ax_i = NULL, *bx_h = NULL, *ax_h = NULL;
    REAL *bz_i = NULL, *az_i = NULL, *bz_h = NULL, *az_h = NULL;
    if (cpml) {
        const int mx = nx > nz ? nx : nz;
        REAL *prof = cpml + 4 * n;
        bx_i = prof; ax_i = prof + mx; bx_h = prof + 2 * mx; ax_h = prof + 3 * mx;
        bz_i = prof + 4 * mx; az_i = prof + 5 * mx; bz_h = prof + 6 * mx; az_h = prof + 7 * mx;
        double tb, ta;
        for (int i = 0; i < nx; ++i) {
            oracle_cpml_coef(i, 0.0, nx, a->sponge, dx, dt, a->pml_r, a->pml_vmax, a->pml_f0, &tb, &ta);
            bx_i[i] = (REAL)tb; ax_i[i] = (REAL)ta;
            oracle_cpml_coef(i, 0.5, nx, a->sponge, dx, dt, a->pml_r, a->pml_vmax, a->pml_f0, &tb, &ta);
            bx_h[i] = (REAL)tb; ax_h[i] = (REAL)ta;
        }
        for (int j = 0; j < nz; ++j) {
            oracle_cpml_coef(j, 0.0, nz, a->sponge, dz, dt, a->pml_r, a->pml_vmax, a->pml_f0, &tb, &ta);
            bz_i[j] = (REAL)tb; az_i[j] = (REAL)ta;
            oracle_cpml_coef(j, 0.5, nz, a->sponge, dz, dt, a->pml_r, a->pml_vmax, a->pml_f0, &tb, &ta);
            bz_h[j] = (REAL)tb; az_h[j] = (REAL)ta;
        }
        if (pml)
            for (size_t c = 0; c < 4 * n; ++c) cpml[c] = pml[c];
    }

    if (state) {
        for (size_t c = 0; c < n; ++c) {
            p2[c] = p1[c] = state[c];
            vx2[c] = vx1[c] = state[n + c];
            vz2[c] = vz1[c] = state[2 * n + c];
        }
    }

    const size_t src = (size_t)a->sx * nz + a->sz;
    int snap_next = 0;
#ifdef _OPENMP
    const int nth = a->nthreads > 0 ? a->nthreads : 1;
#endif

    for (int it = 0; it < a->nt; ++it) {
        const int k = a->k0 + it;
        /* Source before the update, :247-251, in float: ((dt*DEN)*Vp)*signal. */
        p2[src] = p2[src] - dt * (REAL)a->den[src] * (REAL)a->vp[src] * (REAL)a->wavelet[k];

        /* Vx, :255-269 (x-direction stencil, over rows). */
#pragma omp parallel for num_threads(nth) schedule(static)
        for (int i = R; i < nx - R; ++i) {
            for (int j = R; j < nz - R; ++j) {
                const size_t c = (size_t)i * nz + j;
                REAL s = C[0] * (p2[c + nz] - p2[c]);
                for (int m = 2; m <= R; ++m)
                    s = s + C[m - 1] * (p2[c + (size_t)m * nz] - p2[c - (size_t)(m - 1) * nz]);
                if (cpml) {
                    const REAL d = s / (REAL)dx; /* the C-PML terms are float throughout */
                    psi_px[c] = bx_h[i] * psi_px[c] + ax_h[i] * d;
                    s = s + psi_px[c] * (REAL)dx;
                }
                vx1[c] = VEL_UPDATE(vx2[c], a->den[c], dtdx, s);
            }
        }
        /* Vz, :272-286 (z-direction stencil, along the contiguous axis). */
#pragma omp parallel for num_threads(nth) schedule(static)
        for (int i = R; i < nx - R; ++i) {
            for (int j = R; j < nz - R; ++j) {
                const size_t c = (size_t)i * nz + j;
                REAL s = C[0] * (p2[c + 1] - p2[c]);
                for (int m = 2; m <= R; ++m)
                    s = s + C[m - 1] * (p2[c + m] - p2[c - (m - 1)]);
                if (cpml) {
                    const REAL d = s / (REAL)dz;
                    psi_pz[c] = bz_h[j] * psi_pz[c] + az_h[j] * d;
                    s = s + psi_pz[c] * (REAL)dz;
                }
                vz1[c] = VEL_UPDATE(vz2[c], a->den[c], dtdz, s);
            }
        }
        /* P, :289-314, from the new velocities. */
#pragma omp parallel for num_threads(nth) schedule(static)
        for (int i = R; i < nx - R; ++i) {
            for (int j = R; j < nz - R; ++j) {
                const size_t c = (size_t)i * nz + j;
                REAL sx = C[0] * (vx1[c] - vx1[c - nz]);
                for (int m = 2; m <= R; ++m)
                    sx = sx + C[m - 1] * (vx1[c + (size_t)(m - 1) * nz] - vx1[c - (size_t)m * nz]);
                REAL sz = C[0] * (vz1[c] - vz1[c - 1]);
                for (int m = 2; m <= R; ++m)
                    sz = sz + C[m - 1] * (vz1[c + (m - 1)] - vz1[c - m]);
                if (cpml) {
                    psi_vx[c] = bx_i[i] * psi_vx[c] + ax_i[i] * (sx / (REAL)dx);
                    psi_vz[c] = bz_i[j] * psi_vz[c] + az_i[j] * (sz / (REAL)dz);
                    sx = sx + psi_vx[c] * (REAL)dx;
                    sz = sz + psi_vz[c] * (REAL)dz;
                }
                p1[c] = P_UPDATE(p2[c], a->den[c], a->vp[c], dt, dx, dz, sx, sz);
            }
        }
        /* Time-level copy, :317-325, then the sponge on all six arrays, :329-340. */
#pragma omp parallel for num_threads(nth) schedule(static)
        for (size_t c = 0; c < n; ++c) {
            p2[c] = p1[c];
            vx2[c] = vx1[c];
            vz2[c] = vz1[c];
        }
        if (absorb) {
#pragma omp parallel for num_threads(nth) schedule(static)
            for (size_t c = 0; c < n; ++c) {
                const REAL f = (REAL)absorb[c];
                vx1[c] = vx1[c] * f;
                vx2[c] = vx2[c] * f;
                vz1[c] = vz1[c] * f;
                vz2[c] = vz2[c] * f;
                p2[c] = p2[c] * f;
                p1[c] = p1[c] * f;
            }
        }
        /* Snapshots, :343-373 (the reference writes Vx1 at k == 500 and P1 every
         * 10 steps; here any step, all three fields). */
        while (snaps && snap_next < a->nsnap && a->snap_steps[snap_next] == k) {
            REAL *s = snaps + (size_t)snap_next * 3 * n;
            for (size_t c = 0; c < n; ++c) {
                s[c] = p1[c];
                s[n + c] = vx1[c];
                s[2 * n + c] = vz1[c];
            }
            ++snap_next;
        }
        /* Receivers, :376-377: the absorbed new pressure. */
        if (record)
            for (int r = 0; r < a->nrec; ++r)
                record[(size_t)r * a->nt + it] = p1[(size_t)a->rx[r] * nz + a->rz[r]];
    }

    if (state) {
        for (size_t c = 0; c < n; ++c) {
            state[c] = p1[c];
            state[n + c] = vx1[c];
            state[2 * n + c] = vz1[c];
        }
    }
    if (cpml && pml)
        for (size_t c = 0; c < 4 * n; ++c) pml[c] = cpml[c];

    free(p1); free(p2); free(vx1); free(vx2); free(vz1); free(vz2); free(cpml);
    if (absorb && absorb != a->absorb) free(absorb);
    return 0;
fail:
    free(p1); free(p2); free(vx1); free(vx2); free(vz1); free(vz2); free(cpml);
    if (absorb && absorb != a->absorb) free(absorb);
    return -2;
}

#undef VEL_UPDATE
#undef P_UPDATE
