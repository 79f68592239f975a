// Acoustic shot plan shared by the forward, batch and gradient entry points.
#pragma once
#include <vector>

#include "acoustic_kernel.h"
#include "common.h"

namespace sfd {

using StepKernel = void (*)(AcousticStepArgs);
using BatchKernel = void (*)(BatchStepArgs);

StepKernel pick_step_kernel(int radius, int nw, int mode);
BatchKernel pick_batch_kernel(int radius, int nw, int mode);
// Two-launch tiled step (velocity, pressure) of `mode` at tx rows x nw waves
// per tile; false where no such instance exists.  Single plans have every
// shape listed in acoustic.hip for C-PML and sponge, the gradient modes and
// batches the default shape only.
bool pick_tiled(int radius, int tx, int nw, int mode, StepKernel* v, StepKernel* p);
bool pick_tiled_batch(int radius, int tx, int nw, int mode, BatchKernel* v, BatchKernel* p);
// One-launch step of `mode` (C-PML or sponge) at `rows` rows x 64 columns
// per workgroup of nw waves: the tiled step (acoustic_fused.h) for 16 or 32
// rows, the streaming step (acoustic_stream.h) for chunks of kStreamMinRows
// rows or more at 4 or 8 waves; k and b (either may be null) receive the
// single-plan and the batched kernel.  ubf: every updated cell's bf is the
// argument bf_const (uniform density; the streaming step then reads no bf
// field).  False where no such instance exists.
bool pick_fused(int radius, int rows, int nw, int mode, bool ubf, StepKernel* k, BatchKernel* b);
// Launch shapes (rows, waves) the one-launch step is tuned over; the first
// is the default.
constexpr int kFusedShapes[][2] = {{32, 8}, {16, 4}, {16, 8}, {32, 4}};
constexpr int kStreamMinRows = 64;
constexpr int kStreamShapes[][2] = {{128, 4}, {192, 4}, {64, 4}, {128, 8}};
int pad_rows(int radius);
int round_chunk(int ch, int radius);
// Host-side argument checks of an acoustic shot (host receiver arrays).
int validate(sfd_context* ctx, const sfd_acoustic_params* p, const float* vp, const float* wavelet);

}  // namespace sfd

struct sfd_acoustic_plan {
    using StepKernel = sfd::StepKernel;
    using Field = sfd::Field;
    using DeviceBuffer = sfd::DeviceBuffer;
    sfd_context* ctx = nullptr;
    int nx = 0, nz = 0, nt = 0, radius = 0, mode = 0;
    float dx = 0, dz = 0, dt = 0;
    int width = 0;
    int nrec = 0, nslot = 0, nsnap = 0, snap_fields = 0, nsnap_fields = 0;
    std::vector<int> snap_steps;
    float coef[8] = {};
    int pml_xw = 0, pml_zw = 0;
    int sx = 0, sz = 0;
    int pad = 0;  // zero rows above and below the grid in the state buffers
    int nw = 0, chunk = 0;
    int cur = 0;  // ping-pong index holding the current time level

    // One arena holds every field the step kernel addresses (state, model,
    // C-PML strips), so the kernel needs a single buffer descriptor.
    DeviceBuffer arena;
    Field p[2], vx[2], vz[2];
    Field kf, bf;
    Field psi_px[2], psi_vx[2], psi_pz[2], psi_vz[2];
    DeviceBuffer ax, az, pml_x, pml_z;
    DeviceBuffer amp, rec_row, rec_more, rec_buf, rec_slot, record, snaps;
    StepKernel kernel = nullptr;
    // Fast arithmetic (C-PML or sponge) runs the two-launch tiled step
    // instead (acoustic_tiled.h): tile rows and waves, 0 = the march kernel.
    int ttx = 0, tnw = 0;
    StepKernel tvel = nullptr, tpres = nullptr;
    // Or the one-launch tiled step (acoustic_fused.h, ping-pong buffers):
    // tile rows and waves per workgroup, 0 = not selected.
    int ftx = 0, fnw = 0;
    StepKernel fstep = nullptr;
    // The velocity fields hold unabsorbed values (after a tiled sponge step).
    bool v_raw = false;
    // Fast arithmetic with uniform density: bf of every updated cell (the
    // tiled step takes it as an argument); 0 otherwise.
    float bf_const = 0.f;

    // Gradient workspace (plans made by the gradient entry points): adjoint
    // state, gradient accumulator and the divergence of one checkpoint
    // segment live in the arena; checkpoints, residuals and injection weights
    // in their own buffers.
    int grad_period = 0;  // checkpoint interval; 0 = no gradient workspace
    int div_fields = 0;   // divergence fields kept (grad_period, or 1 with boundary saving)
    Field ap[2], avx[2], avz[2], grad, div;
    DeviceBuffer ckpt, res, inj_w, inj_cell, src_adj;
    DeviceBuffer div_mem;  // the divergence ring (pl->div)
    // Boundary saving: offsets (from row 0) of the band cells just outside
    // the reconstructed box and their P, Vx, Vz after every forward step,
    // [nt][3][nband].
    int nband = 0;
    DeviceBuffer band_cell, bands;

    // Row 0 of a padded per-row allocation.
    float* grid(const Field& b) const { return b.as<float>() + (size_t)pad * nz; }
    // Floats between consecutive divergence fields of the gradient workspace.
    size_t div_stride() const { return (size_t)(nx + 2 * pad) * nz; }
    float* grid_period_div() const { return grad_period > 0 ? grid(div) : nullptr; }
};

namespace sfd {

// Model-derived fields shared by the plans of one call on one model (the
// same vp and den arrays, grid, dt, radius and arithmetic): the first plan
// fills it; later plans take its checked maximum velocity and copy its kf and
// bf on the device instead of recomputing and uploading them.
struct PlanModelCache {
    bool valid = false;
    const float *vp = nullptr, *den = nullptr;
    int nx = 0, nz = 0, radius = 0, mode = -1;
    float dt = 0.f, vmax = 0.f, bf_const = 0.f;
    const void *kf = nullptr, *bf = nullptr;  // the first plan's device fields
};

// Creates a plan; grad_period > 0 also reserves the gradient workspace with
// div_fields divergence fields (0 = grad_period of them).
int plan_create(sfd_context* ctx, const sfd_acoustic_params* p, const float* vp, const float* den,
                const float* wavelet, int grad_period, sfd_acoustic_plan** out, int div_fields = 0,
                PlanModelCache* mc = nullptr);
// Step arguments over the state fields (p, vx, vz)[cur] -> [cur ^ 1].
AcousticStepArgs make_args(const sfd_acoustic_plan* pl, const Field* p, const Field* vx, const Field* vz, int cur,
                           int k);
AcousticStepArgs make_args(const sfd_acoustic_plan* pl, int cur, int k);
int enqueue_steps(sfd_acoustic_plan* pl, int k0, int nsteps);
// Turns step arguments into those of the tiled step: in place on the input
// buffers, the tiled tile counts, velocities read as stored (v_raw).
void tiled_args(const sfd_acoustic_plan* pl, AcousticStepArgs& a, bool v_raw);
// Turns step arguments into those of the one-launch tiled step (ping-pong,
// its tile counts, velocities read as stored).
void fused_args(const sfd_acoustic_plan* pl, AcousticStepArgs& a, bool v_raw);
// Enqueues the two launches of one tiled step of a plan.
int launch_tiled(const sfd_acoustic_plan* pl, StepKernel v, StepKernel p, const AcousticStepArgs& a);
// Tiled step shape used unless one is asked for.
constexpr int kTiledTX = 16, kTiledNW = 4;
int zero_state(sfd_acoustic_plan* pl);
// Margin of the box the gradient's reverse reconstruction covers, [m, n - m)
// on both axes: the cells no absorbing term touches (the sponge width; one
// more for C-PML, whose half-node profiles reach a row further), at least
// the stencil radius.
int rev_margin(const sfd_acoustic_plan* pl);
// SFD_ENAN when the plan's gather or current pressure field is not finite.
int check_plan_finite(sfd_acoustic_plan* pl);
// Sponge factor of cell (i, j) as the step kernel applies it.
float sponge_at(const sfd_acoustic_plan* pl, int i, int j);

}  // namespace sfd
