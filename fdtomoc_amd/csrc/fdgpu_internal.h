// fdgpu_internal.h -- declarations shared by the host translation units.
#pragma once

struct fdgpu_ctx;

// Uploads velocities and converts them to slowness on the device; the caller
// holds the context lock.
int fdgpu_model_update_locked(fdgpu_ctx *c, int model, const float *vel);

// Address of the sweep kernel, for occupancy queries.
extern "C" const void *fd_sweep_kernel_ptr();
