/*
 * CPU oracle: discrete adjoint of the acoustic propagator (FP64).
 *
 * The reference has no adjoint; this is the exact transpose, operation by
 * operation, of the forward time step of acoustic_body.h (which restates
 * forward/Acoustic_FD_by_ss/src/Acoustic_1order_FD.cpp:247-377) in double:
 *
 *   P~  = P - amp_k e_s,            amp_k = dt rho_s vp_s w_k     (:247-251)
 *   V'  = [U] (V - b D+ P~),        b = dt / rho                  (:255-286)
 *   P'  = [U] (P~ - K div),  div = D- Vx' + D- Vz', K = rho vp^2 dt (:289-314)
 *   s_k = a (P', V')                sponge factor a               (:329-340)
 *   d_k = S P(s_k)                  receivers                     (:376-377)
 *
 * ([U] = the updated cells [R, N-R)^2; everything else stays 0.)  For a
 * linear functional J = sum_k <seed_k, d_k> the gradient with respect to vp
 * is returned; the L2 misfit gradient is the case seed = d - d_obs.  Every
 * transposed stencil is written as the scatter of its forward gather, so the
 * code is checked by finite differences of the forward without relying on
 * summation-by-parts identities.
 *
 * TEST INFRASTRUCTURE ONLY.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/*
 * C-PML (boundary == 1): the transpose of the oracle's C-PML step
 * (acoustic_body.h; no reference counterpart), in derivative units:
 *
 *   P~   = P - amp_k e_s
 *   d_x  = D+x P~,  Qx' = bh_x Qx + ah_x d_x,  Vx' = Vx - b (d_x + Qx')    (z alike)
 *   t_x  = D-x Vx', Wx' = bi_x Wx + ai_x t_x                              (z alike)
 *   P'   = P~ - K (t_x + Wx' + t_z + Wz')
 *
 * on the updated cells (Q = psi_p*, W = psi_v*; b = 1, a = 0 outside the
 * strips, so the memory fields stay zero there).  The profiles depend on
 * pml_vmax, which the caller fixes: the gradient holds them constant.
 */
static int cpml_gradient(const oracle_acoustic_args *a, const double *seed, double *grad_vp, double *record)
{
    const int nx = a->nx, nz = a->nz, R = a->radius, nt = a->nt;
    if (nx <= 2 * R || nz <= 2 * R || R < 1 || R > 8 || nt < 1 || a->k0 != 0) return -1;
    const size_t n = (size_t)nx * nz;
    const int mx = nx > nz ? nx : nz;
    /* state, its memory fields, the adjoints of both, per-step divergences */
    double *buf = calloc(22 * n + 8 * (size_t)mx, sizeof(double));
    double *divs = calloc((size_t)nt * n, sizeof(double));
    if (!buf || !divs) {
        free(buf);
        free(divs);
        return -2;
    }
    double *p = buf, *vx = p + n, *vz = vx + n, *qx = vz + n, *qz = qx + n, *wx = qz + n, *wz = wx + n;
    double *vxn = wz + n, *vzn = vxn + n;
    double *lp = vzn + n, *lvx = lp + n, *lvz = lvx + n, *lqx = lvz + n, *lqz = lqx + n, *lwx = lqz + n,
           *lwz = lwx + n, *tvx = lwz + n, *tvz = tvx + n, *lpt = tvz + n, *gk = lpt + n;
    double *bxi = gk + n, *axi = bxi + mx, *bxh = axi + mx, *axh = bxh + mx;
    double *bzi = axh + mx, *azi = bzi + mx, *bzh = azi + mx, *azh = bzh + mx;
    for (int i = 0; i < nx; ++i) {
        oracle_cpml_coef(i, 0.0, nx, a->sponge, a->dx, a->dt, a->pml_r, a->pml_vmax, a->pml_f0, &bxi[i], &axi[i]);
        oracle_cpml_coef(i, 0.5, nx, a->sponge, a->dx, a->dt, a->pml_r, a->pml_vmax, a->pml_f0, &bxh[i], &axh[i]);
    }
    for (int j = 0; j < nz; ++j) {
        oracle_cpml_coef(j, 0.0, nz, a->sponge, a->dz, a->dt, a->pml_r, a->pml_vmax, a->pml_f0, &bzi[j], &azi[j]);
        oracle_cpml_coef(j, 0.5, nz, a->sponge, a->dz, a->dt, a->pml_r, a->pml_vmax, a->pml_f0, &bzh[j], &azh[j]);
    }
    double C[8];
    for (int m = 0; m < R; ++m) C[m] = a->coef[m];
    const double dt = a->dt, dx = a->dx, dz = a->dz;
    const size_t src = (size_t)a->sx * nz + a->sz;
#define Bc(c) (dt / (double)a->den[c])
#define Kc(c) ((double)a->den[c] * (double)a->vp[c] * (double)a->vp[c] * dt)

    /* ---- forward, keeping t_x + Wx' + t_z + Wz' of every step ---- */
    for (int k = 0; k < nt; ++k) {
        p[src] -= dt * a->den[src] * (double)a->vp[src] * a->wavelet[k];
        for (int i = R; i < nx - R; ++i)
            for (int j = R; j < nz - R; ++j) {
                const size_t c = (size_t)i * nz + j;
                double ddx = 0, ddz = 0;
                for (int m = 1; m <= R; ++m) {
                    ddx += C[m - 1] / dx * (p[c + (size_t)m * nz] - p[c - (size_t)(m - 1) * nz]);
                    ddz += C[m - 1] / dz * (p[c + m] - p[c - (m - 1)]);
                }
                qx[c] = bxh[i] * qx[c] + axh[i] * ddx;
                qz[c] = bzh[j] * qz[c] + azh[j] * ddz;
                vxn[c] = vx[c] - Bc(c) * (ddx + qx[c]);
                vzn[c] = vz[c] - Bc(c) * (ddz + qz[c]);
            }
        memcpy(vx, vxn, n * sizeof(double));
        memcpy(vz, vzn, n * sizeof(double));
        double *div = divs + (size_t)k * n;
        for (int i = R; i < nx - R; ++i)
            for (int j = R; j < nz - R; ++j) {
                const size_t c = (size_t)i * nz + j;
                double tx = 0, tz = 0;
                for (int m = 1; m <= R; ++m) {
                    tx += C[m - 1] / dx * (vx[c + (size_t)(m - 1) * nz] - vx[c - (size_t)m * nz]);
                    tz += C[m - 1] / dz * (vz[c + (m - 1)] - vz[c - m]);
                }
                wx[c] = bxi[i] * wx[c] + axi[i] * tx;
                wz[c] = bzi[j] * wz[c] + azi[j] * tz;
                div[c] = tx + wx[c] + tz + wz[c];
                p[c] -= Kc(c) * div[c];
            }
        if (record)
            for (int r = 0; r < a->nrec; ++r)
                record[(size_t)r * nt + k] = p[(size_t)a->rx[r] * nz + a->rz[r]];
    }

    /* ---- adjoint, step by step in reverse ---- */
    double gsrc = 0.0;
    for (int k = nt - 1; k >= 0; --k) {
        for (int r = 0; r < a->nrec; ++r)  /* d_k = S P' */
            lp[(size_t)a->rx[r] * nz + a->rz[r]] += seed[(size_t)r * nt + k];
        memcpy(tvx, lvx, n * sizeof(double));
        memcpy(tvz, lvz, n * sizeof(double));
        memset(lpt, 0, n * sizeof(double));
        /* P' = P~ - K (t_x + Wx' + t_z + Wz'), Wx' = bi Wx + ai t_x: identity
         * part, gradient, the memory fields, and the scatter of t_x, t_z. */
        const double *div = divs + (size_t)k * n;
        for (int i = R; i < nx - R; ++i)
            for (int j = R; j < nz - R; ++j) {
                const size_t c = (size_t)i * nz + j;
                lpt[c] += lp[c];
                gk[c] -= lp[c] * div[c];
                const double mu = Kc(c) * lp[c];
                const double wxt = lwx[c] - mu, wzt = lwz[c] - mu;  /* adjoints of Wx', Wz' in total */
                lwx[c] = bxi[i] * wxt;
                lwz[c] = bzi[j] * wzt;
                const double ltx = -mu + axi[i] * wxt, ltz = -mu + azi[j] * wzt;
                for (int m = 1; m <= R; ++m) {
                    tvx[c + (size_t)(m - 1) * nz] += ltx * C[m - 1] / dx;
                    tvx[c - (size_t)m * nz] -= ltx * C[m - 1] / dx;
                    tvz[c + (m - 1)] += ltz * C[m - 1] / dz;
                    tvz[c - m] -= ltz * C[m - 1] / dz;
                }
            }
        /* Vx' = Vx - b (d_x + Qx'), Qx' = bh Qx + ah d_x, d_x = D+x P~. */
        memset(lvx, 0, n * sizeof(double));
        memset(lvz, 0, n * sizeof(double));
        for (int i = R; i < nx - R; ++i)
            for (int j = R; j < nz - R; ++j) {
                const size_t c = (size_t)i * nz + j;
                lvx[c] = tvx[c];
                lvz[c] = tvz[c];
                const double nux = Bc(c) * tvx[c], nuz = Bc(c) * tvz[c];
                const double qxt = lqx[c] - nux, qzt = lqz[c] - nuz;  /* adjoints of Qx', Qz' in total */
                lqx[c] = bxh[i] * qxt;
                lqz[c] = bzh[j] * qzt;
                const double ldx = -nux + axh[i] * qxt, ldz = -nuz + azh[j] * qzt;
                for (int m = 1; m <= R; ++m) {
                    lpt[c + (size_t)m * nz] += ldx * C[m - 1] / dx;
                    lpt[c - (size_t)(m - 1) * nz] -= ldx * C[m - 1] / dx;
                    lpt[c + m] += ldz * C[m - 1] / dz;
                    lpt[c - (m - 1)] -= ldz * C[m - 1] / dz;
                }
            }
        /* P~ = P - dt rho_s vp_s w_k e_s */
        gsrc += lpt[src] * (-dt * a->den[src] * a->wavelet[k]);
        /* P cells outside the updated region are never written (stay zero). */
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < nz; ++j) {
                const size_t c = (size_t)i * nz + j;
                lp[c] = (i >= R && i < nx - R && j >= R && j < nz - R) ? lpt[c] : 0.0;
            }
    }
    for (size_t c = 0; c < n; ++c) grad_vp[c] = gk[c] * 2.0 * (double)a->den[c] * (double)a->vp[c] * dt;
    grad_vp[src] += gsrc;
#undef Bc
#undef Kc
    free(buf);
    free(divs);
    return 0;
}

int oracle_acoustic_gradient_f64(const oracle_acoustic_args *a, const double *seed, double *grad_vp,
                                 double *record)
{
    if (a->boundary == 1) return cpml_gradient(a, seed, grad_vp, record);
    const int nx = a->nx, nz = a->nz, R = a->radius, nt = a->nt;
    if (a->boundary != 0 || nx <= 2 * R || nz <= 2 * R || R < 1 || R > 8 || nt < 1 || a->k0 != 0)
        return -1;
    const size_t n = (size_t)nx * nz;
    double *p = calloc(n, sizeof(double)), *vx = calloc(n, sizeof(double)), *vz = calloc(n, sizeof(double));
    double *vxn = calloc(n, sizeof(double)), *vzn = calloc(n, sizeof(double)), *pn = calloc(n, sizeof(double));
    double *divs = calloc((size_t)nt * n, sizeof(double));
    double *lp = calloc(n, sizeof(double)), *lvx = calloc(n, sizeof(double)), *lvz = calloc(n, sizeof(double));
    double *mp = calloc(n, sizeof(double)), *tvx = calloc(n, sizeof(double)), *tvz = calloc(n, sizeof(double));
    double *lpt = calloc(n, sizeof(double)), *gk = calloc(n, sizeof(double));
    float *absorb = malloc(n * sizeof(float));
    int rc = -2;
    if (!p || !vx || !vz || !vxn || !vzn || !pn || !divs || !lp || !lvx || !lvz || !mp || !tvx || !tvz || !lpt ||
        !gk || !absorb)
        goto done;
    oracle_sponge_table(nx, nz, a->sponge, absorb);

    double C[8];
    for (int m = 0; m < R; ++m) C[m] = a->coef[m];
    const double dt = a->dt, dx = a->dx, dz = a->dz;
    const size_t src = (size_t)a->sx * nz + a->sz;
#define UPD(i, j) ((i) >= R && (i) < nx - R && (j) >= R && (j) < nz - R)
#define Bc(c) (dt / (double)a->den[c])
#define Kc(c) ((double)a->den[c] * (double)a->vp[c] * (double)a->vp[c] * dt)

    /* ---- forward, keeping the divergence of every step ---- */
    for (int k = 0; k < nt; ++k) {
        p[src] -= dt * a->den[src] * (double)a->vp[src] * a->wavelet[k];
        memset(vxn, 0, n * sizeof(double));
        memset(vzn, 0, n * sizeof(double));
        memset(pn, 0, n * sizeof(double));
        for (int i = R; i < nx - R; ++i)
            for (int j = R; j < nz - R; ++j) {
                const size_t c = (size_t)i * nz + j;
                double sx = 0, sz = 0;
                for (int m = 1; m <= R; ++m) {
                    sx += C[m - 1] / dx * (p[c + (size_t)m * nz] - p[c - (size_t)(m - 1) * nz]);
                    sz += C[m - 1] / dz * (p[c + m] - p[c - (m - 1)]);
                }
                vxn[c] = vx[c] - Bc(c) * sx;
                vzn[c] = vz[c] - Bc(c) * sz;
            }
        double *div = divs + (size_t)k * n;
        for (int i = R; i < nx - R; ++i)
            for (int j = R; j < nz - R; ++j) {
                const size_t c = (size_t)i * nz + j;
                double d = 0;
                for (int m = 1; m <= R; ++m) {
                    d += C[m - 1] / dx * (vxn[c + (size_t)(m - 1) * nz] - vxn[c - (size_t)m * nz]);
                    d += C[m - 1] / dz * (vzn[c + (m - 1)] - vzn[c - m]);
                }
                div[c] = d;
                pn[c] = p[c] - Kc(c) * d;
            }
        for (size_t c = 0; c < n; ++c) {
            p[c] = absorb[c] * pn[c];
            vx[c] = absorb[c] * vxn[c];
            vz[c] = absorb[c] * vzn[c];
        }
        if (record)
            for (int r = 0; r < a->nrec; ++r)
                record[(size_t)r * nt + k] = p[(size_t)a->rx[r] * nz + a->rz[r]];
    }

    /* ---- adjoint, step by step in reverse ---- */
    double gsrc = 0.0;
    for (int k = nt - 1; k >= 0; --k) {
        for (int r = 0; r < a->nrec; ++r)  /* d_k = S P(s_k) */
            lp[(size_t)a->rx[r] * nz + a->rz[r]] += seed[(size_t)r * nt + k];
        for (size_t c = 0; c < n; ++c) {   /* s_k = a (P', V') */
            mp[c] = absorb[c] * lp[c];
            tvx[c] = absorb[c] * lvx[c];
            tvz[c] = absorb[c] * lvz[c];
            lpt[c] = 0.0;
        }
        /* P' = [U](P~ - K div): identity part, gradient, and scatter of div. */
        const double *div = divs + (size_t)k * n;
        for (int i = R; i < nx - R; ++i)
            for (int j = R; j < nz - R; ++j) {
                const size_t c = (size_t)i * nz + j;
                lpt[c] += mp[c];
                gk[c] -= mp[c] * div[c];
                const double t = -Kc(c) * mp[c];
                for (int m = 1; m <= R; ++m) {
                    tvx[c + (size_t)(m - 1) * nz] += t * C[m - 1] / dx;
                    tvx[c - (size_t)m * nz] -= t * C[m - 1] / dx;
                    tvz[c + (m - 1)] += t * C[m - 1] / dz;
                    tvz[c - m] -= t * C[m - 1] / dz;
                }
            }
        /* V' = [U](V - b D+ P~): identity part and scatter of D+. */
        for (size_t c = 0; c < n; ++c) {
            lvx[c] = 0.0;
            lvz[c] = 0.0;
        }
        for (int i = R; i < nx - R; ++i)
            for (int j = R; j < nz - R; ++j) {
                const size_t c = (size_t)i * nz + j;
                lvx[c] = tvx[c];
                lvz[c] = tvz[c];
                const double tx = -Bc(c) * tvx[c], tz = -Bc(c) * tvz[c];
                for (int m = 1; m <= R; ++m) {
                    lpt[c + (size_t)m * nz] += tx * C[m - 1] / dx;
                    lpt[c - (size_t)(m - 1) * nz] -= tx * C[m - 1] / dx;
                    lpt[c + m] += tz * C[m - 1] / dz;
                    lpt[c - (m - 1)] -= tz * C[m - 1] / dz;
                }
            }
        /* P~ = P - dt rho_s vp_s w_k e_s */
        gsrc += lpt[src] * (-dt * a->den[src] * a->wavelet[k]);
        memcpy(lp, lpt, n * sizeof(double));
    }
    for (size_t c = 0; c < n; ++c)
        grad_vp[c] = gk[c] * 2.0 * (double)a->den[c] * (double)a->vp[c] * dt;
    grad_vp[src] += gsrc;
    rc = 0;
#undef UPD
#undef Bc
#undef Kc
done:
    free(p); free(vx); free(vz); free(vxn); free(vzn); free(pn); free(divs);
    free(lp); free(lvx); free(lvz); free(mp); free(tvx); free(tvz); free(lpt); free(gk); free(absorb);
    return rc;
}
