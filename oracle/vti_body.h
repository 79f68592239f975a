/*
 * Body of the VTI qP-acoustic oracle, included by vti.c once per arithmetic
 * (REAL = float / double, FN(x) = x_f32 / x_f64).
 *
 * Restates extrapolation_one_step_1 / _2
 * (de-multiple/convolution_srm/vti_qa_2d_c3_demo1/VTI_multishot_qacoustic_fd_subroutine_c3.f90:432-1258)
 * and the time loop of single_shot_modelling (VTI_multishot_qacoustic_fd_main_c3.f90:429-520).
 * Fields are [nx][nz] here (z fastest); the reference's u(i, j) is column
 * i = x, row j = z.  Every expression keeps the reference's operand order
 * (Fortran evaluates a*b*c as (a*b)*c and ** 2 as a product).
 */

typedef struct {
    int nx, nz, nl, nr, nu, nd, nxs, nzs;
    REAL dx, dz, dt, lnr;
    REAL c2[5], c1[5];
    const float *vz, *eps, *del;
} FN(vti_ctx);

/* Halo-padded index: 5 zero cells on every side, as the reference's
 * u(-4:nx+5, -4:nz+5). */
#define VH 5
#define VIX(c, i, j) ((size_t)((i) + VH) * (size_t)((c)->nz + 2 * VH) + (size_t)((j) + VH))

/* sum_m c_m (f(+m) + f(-m) - 2 f(0)) / h^2 along x (s = 0) or z (s = 1):
 * the reference's sum_x_u_d2 / sum_dx_2 (sub:516-527, 592-601). */
static REAL FN(vti_d2)(const FN(vti_ctx) * c, const REAL *f, int i, int j, int axis, const REAL *cf, REAL h)
{
    REAL s = 0;
    const REAL f0 = f[VIX(c, i, j)];
    for (int m = 1; m <= 5; ++m) {
        const REAL a = axis == 0 ? f[VIX(c, i + m, j)] : f[VIX(c, i, j + m)];
        const REAL b = axis == 0 ? f[VIX(c, i - m, j)] : f[VIX(c, i, j - m)];
        const REAL t = cf[m - 1] * (a + b - 2 * f0);
        s = m == 1 ? t : s + t;
    }
    return s / (h * h);
}

/* sum_m c_m (f(+m) - f(-m)) / h: the reference's sum_dx_1 (sub:566-578). */
static REAL FN(vti_d1)(const FN(vti_ctx) * c, const REAL *f, int i, int j, int axis, const REAL *cf, REAL h)
{
    REAL s = 0;
    for (int m = 1; m <= 5; ++m) {
        const REAL a = axis == 0 ? f[VIX(c, i + m, j)] : f[VIX(c, i, j + m)];
        const REAL b = axis == 0 ? f[VIX(c, i - m, j)] : f[VIX(c, i, j - m)];
        const REAL t = cf[m - 1] * (a - b);
        s = m == 1 ? t : s + t;
    }
    return s / h;
}

/* The split-PML recursions shared by both routines (sub:651-663, 1141-1153):
 *   P1 <- (P1c (2 h^2 + 2 d dt h^2 - d^2 h^2 dt^2) - h^2 P1o + k dt^2 (S2 h^2)) / (h^2 + 2 d dt h^2)
 *   Pt <- (Ptc (2 h + 2 d dt h - dt^2 h d^2) - h Pto - k dt^2 d' (S1 h)) / (h + 2 d dt h)
 *   P2 <- P2c (1 - d dt) + Pt dt
 * with slots [old, cur] of P1 and Pt; the old slot takes the new value. */
static REAL FN(vti_pml)(REAL *p1o, REAL p1c, REAL *pto, REAL ptc, REAL *p2, REAL k, REAL S2, REAL S1, REAL d,
                        REAL dd, REAL h, REAL dt)
{
    const REAL h2 = h * h, dt2 = dt * dt;
    const REAL n1 = p1c * (2 * h2 + 2 * d * dt * h2 - d * d * h2 * dt2) - h2 * *p1o + k * dt2 * (S2 * h2);
    *p1o = n1 / (h2 + 2 * d * dt * h2);
    const REAL nt = ptc * (2 * h + 2 * d * dt * h - dt2 * h * (d * d)) - h * *pto - k * dt2 * dd * (S1 * h);
    *pto = nt / (h + 2 * d * dt * h);
    *p2 = *p2 * (1 - d * dt) + *pto * dt;
    return *p1o + *p2;
}

/* Damping profile at signed distance a (cells x h) into a strip of L = n h:
 * d = 3 v / (2 L) (a / L)^2 ln(1/R), d' = 3 v / (2 L) (2 a / L^2) ln(1/R)
 * (sub:645-649). */
static void FN(vti_damp)(REAL v, REAL a, REAL L, REAL lnr, REAL *d, REAL *dd)
{
    const REAL g = 3 * v / (2 * L);
    *d = g * ((a / L) * (a / L)) * lnr;
    *dd = g * (2 * a / (L * L)) * lnr;
}

/* One time step: u1/q1 (old level) <- new level from u2/q2 (current level,
 * the source already added).  Parts: x strips [12][nxs] = U1 old/cur, Ut
 * old/cur, U2, U3 old/cur, Q1 old/cur, Qt old/cur, Q2; z strips [7][nzs] =
 * the U parts.  `par` selects which slot of a two-level part is old. */
static void FN(vti_step)(const FN(vti_ctx) * c, REAL *u1, const REAL *u2, REAL *q1, const REAL *q2, REAL *xs,
                         REAL *zs, int par, int nthreads)
{
    const int nx = c->nx, nz = c->nz;
    const REAL dx = c->dx, dz = c->dz, dt = c->dt;
    const int o = par, k = 1 - par;
    (void)nthreads;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < nz; ++j) {
            const size_t m = (size_t)i * nz + j, h = VIX(c, i, j);
            const REAL vz = c->vz[m], ep = c->eps[m], de = c->del[m];
            const int inx = i >= c->nl && i < nx - c->nr;
            const int inz = j >= c->nu && j < nz - c->nd;
            /* ---------------------------------------------------- p (step_1) */
            if (inx && inz) {
                const REAL sxu = FN(vti_d2)(c, u2, i, j, 0, c->c2, dx);
                const REAL szu = FN(vti_d2)(c, u2, i, j, 1, c->c2, dz);
                const REAL sxq = FN(vti_d2)(c, q2, i, j, 0, c->c2, dx);
                u1[h] = 2 * u2[h] - u1[h] + dt * dt * (vz * vz) * ((1 + 2 * de) * (sxu + sxq) + szu);
            } else if (!inx) {
                const REAL vn = vz * (REAL)sqrt(1 + (REAL)2 * de), vx = vz * (REAL)sqrt(1 + (REAL)2 * ep);
                const REAL S1 = FN(vti_d1)(c, u2, i, j, 0, c->c1, dx) + FN(vti_d1)(c, q2, i, j, 0, c->c1, dx);
                const REAL S2 = FN(vti_d2)(c, u2, i, j, 0, c->c2, dx) + FN(vti_d2)(c, q2, i, j, 0, c->c2, dx);
                const REAL szu = FN(vti_d2)(c, u2, i, j, 1, c->c2, dz);
                const int left = i < c->nl;
                const REAL a = left ? (REAL)(i - c->nl) * dx : (REAL)(i + 1 - (nx - c->nr)) * dx;
                const REAL L = (REAL)(left ? c->nl : c->nr) * dx;
                const size_t s = (size_t)(left ? i : c->nl + i - (nx - c->nr)) * nz + j, n = c->nxs;
                REAL d, dd;
                FN(vti_damp)(vx, a, L, c->lnr, &d, &dd);
                const REAL ux = FN(vti_pml)(&xs[o * n + s], xs[k * n + s], &xs[(2 + o) * n + s], xs[(2 + k) * n + s],
                                            &xs[4 * n + s], vn * vn, S2, S1, d, dd, dx, dt);
                REAL *u3o = &xs[(5 + o) * n + s];
                *u3o = 2 * xs[(5 + k) * n + s] - *u3o + vz * vz * (dt * dt) * (szu * (dz * dz)) / (dz * dz);
                u1[h] = ux + *u3o;
            } else {
                const REAL vn = vz * (REAL)sqrt(1 + (REAL)2 * de);
                const REAL S1 = FN(vti_d1)(c, u2, i, j, 1, c->c1, dx); /* the reference divides by dx (sub:796) */
                const REAL sxu = FN(vti_d2)(c, u2, i, j, 0, c->c2, dx);
                const REAL szu = FN(vti_d2)(c, u2, i, j, 1, c->c2, dz);
                const REAL sxq = FN(vti_d2)(c, q2, i, j, 0, c->c2, dx);
                const int top = j < c->nu;
                const REAL a = top ? (REAL)(j - c->nu) * dz : (REAL)(j + 1 - (nz - c->nd)) * dz;
                const REAL L = (REAL)(top ? c->nu : c->nd) * dz;
                const size_t s = (size_t)(i - c->nl) * (c->nu + c->nd) + (top ? j : c->nu + j - (nz - c->nd));
                const size_t n = c->nzs;
                REAL d, dd;
                FN(vti_damp)(vz, a, L, c->lnr, &d, &dd);
                const REAL uz = FN(vti_pml)(&zs[o * n + s], zs[k * n + s], &zs[(2 + o) * n + s], zs[(2 + k) * n + s],
                                            &zs[4 * n + s], vz * vz, szu, S1, d, dd, dz, dt);
                REAL *u3o = &zs[(5 + o) * n + s];
                *u3o = 2 * zs[(5 + k) * n + s] - *u3o + vn * vn * (dt * dt) * (sxu + sxq);
                u1[h] = uz + *u3o;
            }
            /* ---------------------------------------------------- q (step_2) */
            if (inx) {
                const REAL sxu = FN(vti_d2)(c, u2, i, j, 0, c->c2, dx);
                const REAL sxq = FN(vti_d2)(c, q2, i, j, 0, c->c2, dx);
                q1[h] = 2 * q2[h] - q1[h] + dt * dt * (vz * vz) * (2 * (ep - de) * (sxq + sxu));
            } else {
                const REAL vn = vz * (REAL)sqrt(1 + (REAL)2 * de), vx = vz * (REAL)sqrt(1 + (REAL)2 * ep);
                const REAL S1 = FN(vti_d1)(c, q2, i, j, 0, c->c1, dx) + FN(vti_d1)(c, u2, i, j, 0, c->c1, dx);
                const REAL S2 = FN(vti_d2)(c, q2, i, j, 0, c->c2, dx) + FN(vti_d2)(c, u2, i, j, 0, c->c2, dx);
                const int left = i < c->nl;
                const REAL a = left ? (REAL)(i - c->nl) * dx : (REAL)(i + 1 - (nx - c->nr)) * dx;
                const REAL L = (REAL)(left ? c->nl : c->nr) * dx;
                const size_t s = (size_t)(left ? i : c->nl + i - (nx - c->nr)) * nz + j, n = c->nxs;
                REAL d, dd;
                FN(vti_damp)(vx, a, L, c->lnr, &d, &dd);
                q1[h] = FN(vti_pml)(&xs[(7 + o) * n + s], xs[(7 + k) * n + s], &xs[(9 + o) * n + s],
                                    xs[(9 + k) * n + s], &xs[11 * n + s], vx * vx - vn * vn, S2, S1, d, dd, dx, dt);
            }
        }
    }
}

int FN(oracle_vti)(const oracle_vti_args *a, REAL *record, REAL *state)
{
    if (!a || a->nx < 1 || a->nz < 1 || a->nsteps < 0 || a->nl < 0 || a->nr < 0 || a->nu < 0 || a->nd < 0 ||
        a->nl + a->nr > a->nx || a->nu + a->nd > a->nz || !a->vz || !a->eps || !a->del || !a->c2 || !a->c1)
        return -1;
    if (a->nwav > 0 && (!a->wavelet || a->sx < 0 || a->sx >= a->nx || a->sz < 0 || a->sz >= a->nz)) return -1;
    for (int r = 0; r < a->nrec; ++r)
        if (a->rx[r] < 0 || a->rx[r] >= a->nx || a->rz[r] < 0 || a->rz[r] >= a->nz) return -1;
    FN(vti_ctx) c;
    c.nx = a->nx;
    c.nz = a->nz;
    c.nl = a->nl;
    c.nr = a->nr;
    c.nu = a->nu;
    c.nd = a->nd;
    c.nxs = (a->nl + a->nr) * a->nz;
    c.nzs = (a->nx - a->nl - a->nr) * (a->nu + a->nd);
    c.dx = a->dx;
    c.dz = a->dz;
    c.dt = a->dt;
    const REAL R = a->r > 0 ? a->r : (REAL)1e-3f;
    c.lnr = (REAL)log(1 / R);
    for (int m = 0; m < 5; ++m) {
        c.c2[m] = a->c2[m];
        c.c1[m] = a->c1[m];
    }
    c.vz = a->vz;
    c.eps = a->eps;
    c.del = a->del;
    const size_t nh = (size_t)(a->nx + 2 * VH) * (a->nz + 2 * VH), ncell = (size_t)a->nx * a->nz;
    REAL *f = calloc(4 * nh, sizeof(REAL));
    REAL *xs = calloc(12 * (size_t)c.nxs + 1, sizeof(REAL)), *zs = calloc(7 * (size_t)c.nzs + 1, sizeof(REAL));
    if (!f || !xs || !zs) {
        free(f);
        free(xs);
        free(zs);
        return -2;
    }
    REAL *u[2] = {f, f + nh}, *q[2] = {f + 2 * nh, f + 3 * nh}; /* [old, cur] */
    if (record) memset(record, 0, sizeof(REAL) * (size_t)a->nrec * (size_t)(a->nt > 0 ? a->nt : 0));
    int cur = 1;
    for (int step = 0; step < a->nsteps; ++step) {
        /* the source before the step, into the current p (main:433-437) */
        if (step < a->nwav) u[cur][VIX(&c, a->sx, a->sz)] += (REAL)a->wavelet[step];
        const int old = 1 - cur;
        FN(vti_step)(&c, u[old], u[cur], q[old], q[cur], xs, zs, step & 1, a->nthreads > 0 ? a->nthreads : 1);
        cur = old;
        const int smp = step - a->delay;
        if (record && smp >= 0 && smp < a->nt)
            for (int r = 0; r < a->nrec; ++r) record[(size_t)r * a->nt + smp] = u[cur][VIX(&c, a->rx[r], a->rz[r])];
    }
    if (state)
        for (size_t m = 0; m < ncell; ++m) {
            const int i = (int)(m / a->nz), j = (int)(m % a->nz);
            state[m] = u[cur][VIX(&c, i, j)];
            state[ncell + m] = q[cur][VIX(&c, i, j)];
        }
    free(f);
    free(xs);
    free(zs);
    return 0;
}

#undef VIX
#undef VH
