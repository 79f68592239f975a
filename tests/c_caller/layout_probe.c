/* Size and field offsets of sfd_visco_params as the C compiler lays it out,
 * one "name offset" line per field, for comparison with the Fortran bind(C)
 * type of seisfd_visco.f90. */
#include <stddef.h>
#include <stdio.h>

#include "seisfd.h"

#define F(x) printf(#x " %zu\n", offsetof(sfd_visco_params, x))

int main(void) {
    printf("size %zu\n", sizeof(sfd_visco_params));
    F(nx); F(nz); F(nt); F(delay); F(dx); F(dz); F(dt); F(f0); F(radius); F(coef); F(nmech);
    F(q_scale); F(f_low); F(f_high); F(sponge_xlo); F(sponge_xhi); F(sponge_zlo); F(sponge_zhi);
    F(sponge_r); F(sx); F(sz); F(src_sigma); F(nrec); F(rx); F(rz); F(nsnap); F(snap_steps);
    F(snap_fields); F(qs);
    return 0;
}
