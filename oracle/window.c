/*
 * CPU oracle: aperture window of one shot, restating get_currshot_parameters
 * (forward/Q_forward/Q_multishot_fd_subroutine_fsbc_zy_rotated.f90:288-373)
 * in its own order: copy the window columns (:315-327), pad the top rows
 * (:331-340) and bottom rows (:343-352) for every column, then the left
 * (:355-366) and right (:369-380) columns.  Fortran indices are translated to
 * 0-based C with the reference's own column mapping, i_ori = i + ixmine.
 *
 * TEST INFRASTRUCTURE ONLY.
 */
#include <stddef.h>

#include "oracle.h"

/* model [nx][nz] (the reference's vp(nz,nx) column-major is the same bytes);
 * ixmine as the reference computes it (1-based); out [nxea][nzea]. */
void oracle_window(const float *model, int nx, int nz, int ixmine, int nxe, int npxn, int npxp,
                   int npzn, int npzp, float *out)
{
    const int nxea = nxe + npxn + npxp, nzea = nz + npzn + npzp;
    /* Window columns i = 1..nxe take model column i_ori = i + ixmine, clamped. */
    for (int i = 1; i <= nxe; ++i) {
        int i_ori = i + ixmine;
        if (i_ori < 1) i_ori = 1;
        if (i_ori > nx) i_ori = nx;
        for (int k = 1; k <= nz; ++k)
            out[(size_t)(i + npxn - 1) * nzea + (npzn + k - 1)] = model[(size_t)(i_ori - 1) * nz + (k - 1)];
    }
    /* Top rows k = 1..npzn copy row npzn+1; bottom rows copy row npzn+nz. */
    for (int k = 1; k <= npzn; ++k)
        for (int i = 1; i <= nxea; ++i)
            out[(size_t)(i - 1) * nzea + (k - 1)] = out[(size_t)(i - 1) * nzea + npzn];
    for (int k = npzn + nz + 1; k <= nzea; ++k)
        for (int i = 1; i <= nxea; ++i)
            out[(size_t)(i - 1) * nzea + (k - 1)] = out[(size_t)(i - 1) * nzea + (npzn + nz - 1)];
    /* Left columns copy column npxn+1, right columns copy column npxn+nxe. */
    for (int k = 1; k <= nzea; ++k) {
        for (int i = 1; i <= npxn; ++i)
            out[(size_t)(i - 1) * nzea + (k - 1)] = out[(size_t)npxn * nzea + (k - 1)];
        for (int i = npxn + nxe + 1; i <= nxea; ++i)
            out[(size_t)(i - 1) * nzea + (k - 1)] = out[(size_t)(npxn + nxe - 1) * nzea + (k - 1)];
    }
}
