// The acoustic reference program as a maintainer would port it to libseisfd:
// the constant block of forward/Acoustic_FD_by_ss/src/Acoustic_1order_FD.cpp
// (:28-45, :137-216) as data, its time loop (:237-380) replaced by one
// sfd_acoustic_forward call in the reference arithmetic, and its two output
// files (Snap.dat, :343-355; record.dat, :385-393) written the same way, into
// the directory given as the first argument.  Built against include/seisfd.h
// and linked to libseisfd.so (tests/c_caller/Makefile).
#include <cmath>
#include <cstdio>
#include <vector>

#include "seisfd.h"

#define PI 3.1415926

int main(int argc, char** argv) {
    const char* out = argc > 1 ? argv[1] : ".";
    const int NX = 100, NZ = 100, PML = 20, Nx = NX + 2 * PML, Nz = NZ + 2 * PML, Nt = 2000;
    const float dx = 10.0, dz = 10.0, dt = 0.0005;
    const float coef[8] = {1.2340911f, -1.0664985e-1f, 2.3036367e-2f, -5.3423856e-3f,
                           1.0772712e-3f, -1.6641888e-4f, 1.7021711e-5f, -8.5234642e-7f};
    // Homogeneous model (the reference's Vp = 2000 m/s, DEN = 2400 kg/m^3).
    std::vector<float> vp((size_t)Nx * Nz, 2000.0f), den((size_t)Nx * Nz, 2400.0f);
    // 28 Hz Ricker of amplitude 50 delayed by one period, evaluated in double
    // through the double PI as the reference's expression does, stored float.
    std::vector<float> signal(Nt);
    const float freq = 28;
    for (int i = 0; i < Nt; i++) {
        const float t = dt * i;
        const float t1 = 1.0 / freq;
        signal[i] = 50 * (1 - 2 * PI * PI * freq * freq * (t - t1) * (t - t1)) *
                    exp(-PI * PI * freq * freq * (t - t1) * (t - t1));
    }
    const int nx_location = Nx / 2, nz_location = Nz / 2;
    // record[i - PML][k] = P1[nx_location][i] (:376-377): a column of receivers.
    std::vector<int> rx(NX, nx_location), rz(NX);
    for (int i = 0; i < NX; i++) rz[i] = PML + i;
    const int snap_step = 500;

    sfd_acoustic_params p = {};
    p.nx = Nx;
    p.nz = Nz;
    p.nt = Nt;
    p.dx = dx;
    p.dz = dz;
    p.dt = dt;
    p.radius = 8;
    p.coef = coef;
    p.boundary = SFD_BOUNDARY_SPONGE;
    p.width = PML;
    p.sx = nx_location;
    p.sz = nz_location;
    p.nrec = NX;
    p.rx = rx.data();
    p.rz = rz.data();
    p.nsnap = 1;
    p.snap_steps = &snap_step;
    p.snap_fields = SFD_FIELD_VX;
    p.arith = SFD_ARITH_REFERENCE;

    std::vector<float> rec((size_t)NX * Nt), snap((size_t)Nx * Nz);
    sfd_context* ctx = nullptr;
    int rc = sfd_context_create(0, &ctx);
    if (rc != SFD_OK) {
        fprintf(stderr, "seisfd: no device (%d)\n", rc);
        return 1;
    }
    rc = sfd_acoustic_forward(ctx, &p, vp.data(), den.data(), signal.data(), rec.data(), snap.data());
    if (rc != SFD_OK) {
        fprintf(stderr, "seisfd: %s\n", sfd_last_error(ctx));
        sfd_context_destroy(ctx);
        return 1;
    }
    sfd_context_destroy(ctx);

    char path[4096];
    snprintf(path, sizeof path, "%s/Snap.dat", out);
    FILE* fs = fopen(path, "wb");
    if (!fs) return 1;
    for (int ii = PML; ii < Nx - PML; ii++)
        for (int jj = PML; jj < Nz - PML; jj++) fwrite(&snap[(size_t)ii * Nz + jj], sizeof(float), 1, fs);
    fclose(fs);
    snprintf(path, sizeof path, "%s/record.dat", out);
    FILE* fr = fopen(path, "wb");
    if (!fr) return 1;
    fwrite(rec.data(), sizeof(float), rec.size(), fr);  // [NX][Nt], trace after trace
    fclose(fr);
    printf("End of Calculation!!\n");
    return 0;
}
