/*
 * fwi3d.h -- C ABI of libfwi3d_hip.so, the MI355X (gfx950) wave-propagation
 * library behind FWI3D's Modeling / Migration / Inversion classes.
 *
 * The reference (phbastosa/FWI3D) has no FFI: its classes launch CUDA kernels
 * directly.  Every entry point below replaces one of those launch sites or
 * the host loop around it; the comment on each names the reference code it
 * stands in for.  Everything is plain C: opaque context, raw pointers, sizes.
 *
 * Conventions
 *  - Every call returns an int status: FWI_OK (0) or a negative FWI_ERR_*.
 *    fwi_last_error() returns a human-readable message for the last failure
 *    on a context (or of fwi_create when ctx is NULL).  No C++ exception
 *    crosses this boundary.
 *  - Host buffers are owned by the caller.  Calls that take or return host
 *    buffers are synchronous on return; volumes live in device memory (HBM)
 *    between calls.
 *  - Host volumes use the reference layout  c = i + j*nzz + k*nxx*nzz
 *    (i = z fastest, j = x, k = y; src/modeling/modeling.cu:190).
 *    Seismograms use the reference file layout  t + r*nt
 *    (src/modeling/modeling.cu:392).
 *  - A context is used from one host thread at a time.  Contexts on
 *    different devices may run concurrently.
 */
#ifndef FWI3D_H
#define FWI3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FWI_OK 0
#define FWI_ERR_ARG (-1)       /* bad argument (NULL pointer, size, enum) */
#define FWI_ERR_HIP (-2)       /* a HIP runtime call failed */
#define FWI_ERR_CFL (-3)       /* model violates v*dt/dh <= 0.452856 */
#define FWI_ERR_STATE (-4)     /* call order: e.g. backward before forward */
#define FWI_ERR_RANGE (-5)     /* source/receiver outside the updated grid; padded x-y plane >= 2 GiB */
#define FWI_ERR_NONFINITE (-6) /* NaN/Inf found in a model or wavefield */
#define FWI_ERR_NODEV (-7)     /* no HIP device */

/* boundary treatment of a forward pass */
#define FWI_BOUNDARY_RBC 0 /* no damping: random-boundary model, reversible */
#define FWI_BOUNDARY_ABC 1 /* Cerjan damping (modeling.cu:374-380) */

/* arithmetic of the stencil */
#define FWI_ARITH_REFERENCE 0 /* per-axis derivative summed in FP64, as the
                                 reference's double FD literals do
                                 (modeling.cuh:9-13); bit-exact with the oracle */
#define FWI_ARITH_FP32 1      /* all-FP32 FMA stencil */

/* volumes for fwi_get_volume / fwi_set_volume */
#define FWI_VOL_MODEL 0     /* vp (padded) */
#define FWI_VOL_FIELD 1     /* source-side current field (d_P) */
#define FWI_VOL_FIELD_OLD 2 /* source-side previous field (d_Pold) */
#define FWI_VOL_RECV 3      /* receiver-side current field (d_Pr) */
#define FWI_VOL_RECV_OLD 4  /* receiver-side previous field (d_Prold) */
#define FWI_VOL_IMAGE 5     /* RTM image accumulator (d_image) */
#define FWI_VOL_ILLUM 6     /* illumination accumulator (d_sumPs) */
#define FWI_VOL_GRAD 7      /* FWI gradient accumulator (d_gradient_rbc) */

/* Padded grid.  nxx = nx + 2*nb etc. (modeling.cu:146-150). */
typedef struct fwi_grid {
    int nxx, nyy, nzz; /* padded samples in x, y, z */
    int nb;            /* boundary cells per side */
    float dh;          /* grid spacing [m] */
    float dt;          /* time step [s] */
    int nt;            /* recorded samples (time_samples) */
    int tlag;          /* wavelet lag in steps; a pass runs nt + tlag steps */
} fwi_grid;

typedef struct fwi_ctx fwi_ctx;

/* Create a context on `device` and allocate the padded model and the two
 * source-side fields.  Stands in for Modeling::set_properties' device
 * allocations (modeling.cu:142-179).  Receiver-side fields, image,
 * illumination and gradient are allocated on first use.  Padded volumes of
 * 2 GiB and more run step-kernel instances that address each z-plane with
 * its own 64-bit base (one shot per block, image and illumination updated
 * every step); one padded x-y plane must stay below 2 GiB.  Default launch
 * shapes: 64 x 32 tiles for the forward passes and 64 x 16 for the backward
 * ones, with the fewest z-chunks that give about one block per CU
 * (fwi_autotune times the alternatives). */
int fwi_create(const fwi_grid* grid, int device, fwi_ctx** out);
int fwi_destroy(fwi_ctx* ctx);
const char* fwi_last_error(const fwi_ctx* ctx);

/* FWI_ARITH_REFERENCE (default) or FWI_ARITH_FP32. */
int fwi_set_arith(fwi_ctx* ctx, int arith);

/* Optional corrections of the backward passes (bit mask; 0, the default,
 * reproduces the reference).  No reference counterpart: the three flags
 * switch off the reference behaviours of inversion.cu:466 (gradient without
 * d2/dy2), migration.cu:195 / inversion.cu:461 (reconstruction keeps the
 * injected source) and migration.cu:195-200 (the reconstructed source field
 * is paired with the receiver field one step later). */
#define FWI_PHYS_GRAD_DYY 1          /* FWI gradient with d2/dy2 */
#define FWI_PHYS_SUBTRACT_SOURCE 2   /* remove the source term during reconstruction */
#define FWI_PHYS_ALIGNED_IMAGING 4   /* image/illumination from the source field at the receiver field's time */
int fwi_set_physics(fwi_ctx* ctx, int flags);

/* Backward passes update the image every other time step and the
 * illumination every third, adding the held terms in step order (enable =
 * 1, the default; results are bit-identical to per-step updates), or both
 * every step (0).  Terms still pending when a pass or a segment of
 * backward_steps ends are added by one extra launch.  With aligned imaging,
 * and in batched blocks of 1024 threads with two or more rows per thread,
 * both are updated every other step; batched passes with corrections, and
 * batched FWI passes in blocks of more than 512 threads, update every step;
 * the FWI gradient term is always added per step, only its illumination
 * deferred.  No reference counterpart: the
 * reference's cross_correlation (migration.cu:181-202) and build_gradient
 * (inversion.cu:400-468) update them in every step. */
int fwi_set_deferred_imaging(fwi_ctx* ctx, int enable);

/* Padded velocity model, host layout (cudaMemcpy of d_Vp, modeling.cu:176).
 * Rejects NaN/Inf and CFL violations. */
int fwi_set_model(fwi_ctx* ctx, const float* vp_padded);

/* Interior model (nz*nx*ny, host layout) padded on the device by edge
 * replication (Modeling::expand_boundary, modeling.cu:181-228). */
int fwi_set_model_interior(fwi_ctx* ctx, const float* vp_interior);

/* Seeded random boundary on the current model, in place on the device:
 * v = v_face*(1 + amp*ramp*u), u ~ U[-1,1) from a counter hash of
 * (seed, shot, cell), ramp = 1 - d/nb, clamped to vclamp when vclamp > 0.
 * Replaces Modeling::set_random_boundary (modeling.cu:558-666), whose curand
 * and Poisson-disk noise are not reproducible. */
int fwi_random_boundary(fwi_ctx* ctx, uint64_t seed, int shot, float amp, float vclamp);

/* Reference-style random boundary (random_boundary_bg + random_boundary_gp,
 * modeling.cu:558-666) in one fused launch instead of one launch per blob,
 * deterministic for (seed, shot): each boundary cell starts from the nearest
 * interior cell and gets per-axis noise (1 - d/nb) * var_vp * n, n ~ N(0,1)
 * (Irwin-Hall of 12 hashed uniforms), then Gaussian blobs at Poisson-disk
 * points (minimum distance `ratio` metres) outside the inner box, radius
 * U[0.2, 1) ratio, amplitude +-U[0.5, 1) var_vp, each truncated at 4 radii,
 * with the clamp to [vmin - var_vp, vmax + var_vp] (extremes of the interior
 * model) after every blob, then min(v, vclamp).  Only the nb-cell boundary is
 * written, so a repeated call starts afresh.  Stands in for
 * Modeling::set_random_boundary(Vp, ratio, varVp) (modeling.cu:558). */
int fwi_random_boundary_blobs(fwi_ctx* ctx, uint64_t seed, int shot, float ratio, float var_vp, float vclamp);

/* The blob list of fwi_random_boundary_blobs for a padded grid (host only,
 * no device needed): *count rows of (xc, yc, zc, r, signed amplitude) in
 * metres and m/s, written to out when cap >= *count. */
int fwi_rbc_blobs(uint64_t seed, int shot, int nxx, int nyy, int nzz, int nb, float dh, float ratio, float var_vp,
                  float* out, int cap, int* count);

/* Source wavelet, nt samples (cudaMemcpy of d_wavelet, modeling.cu:54-56). */
int fwi_set_wavelet(fwi_ctx* ctx, const float* wavelet);

/* 1-D Cerjan damper b1d[nb] (modeling.cu:96-99); the 2-D and 3-D dampers of
 * get_boundary_damper (modeling.cu:395-520) are derived from it on the fly
 * with the same float operations as the reference's b2d/b3d tables. */
int fwi_set_dampers(fwi_ctx* ctx, const float* b1d);

/* Source and receiver padded indices (Modeling::initialization,
 * modeling.cu:264-300).  Positions must lie in the updated interior
 * 4 <= idx < n-4. */
int fwi_set_source(fwi_ctx* ctx, int sIdx, int sIdy, int sIdz);
int fwi_set_receivers(fwi_ctx* ctx, int nrec, const int* rIdx, const int* rIdy, const int* rIdz);

/* Full forward pass, nt + tlag steps from zero fields, recording the
 * seismogram on the device (Modeling::forward_solver with compute_pressure
 * and compute_seismogram, modeling.cu:309-393). */
int fwi_forward(fwi_ctx* ctx, int boundary);

/* Steps [n_begin, n_end) of the forward loop from the current fields
 * (no reset); for step-wise checks. */
int fwi_forward_steps(fwi_ctx* ctx, int boundary, int n_begin, int n_end);

/* Recorded seismogram, nt*nrec floats, layout t + r*nt (cudaMemcpy of
 * d_seismogram, modeling.cu:304).  Sample 0 of every trace is 0. */
int fwi_get_seismogram(fwi_ctx* ctx, float* out);

/* Upload the data injected by the backward pass (nt*nrec, t + r*nt):
 * Migration::set_seismic_source (migration.cu:83-88). */
int fwi_set_backward_data(fwi_ctx* ctx, const float* data);

/* Full RTM backward pass (Migration::backward_propagation with
 * inject_seismogram and cross_correlation, migration.cu:61-81,119-202).
 * Zeroes the receiver fields, injects the uploaded data time-reversed and
 * accumulates image and illumination (not reset between shots). */
int fwi_rtm_backward(fwi_ctx* ctx);

/* Full FWI backward pass (Inversion::backward_propagation with
 * inject_residuo and build_gradient, inversion.cu:235-253,385-468); uses the
 * uploaded data (the residual) and accumulates gradient and illumination. */
int fwi_grad_backward(fwi_ctx* ctx);

/* Backward steps [b_begin, b_end) from the current fields; kind 0 = RTM,
 * 1 = FWI.  b_begin == 0 zeroes nothing: call fwi_zero_volume first. */
int fwi_backward_steps(fwi_ctx* ctx, int kind, int b_begin, int b_end);

/* Zero one volume (cudaMemset in the reference, e.g. migration.cu:81-82). */
int fwi_zero_volume(fwi_ctx* ctx, int which);

/* Copy a padded volume to / from the host (host layout). */
int fwi_get_volume(fwi_ctx* ctx, int which, float* out_padded);
int fwi_set_volume(fwi_ctx* ctx, int which, const float* in_padded);

/* Migration::export_seismic arithmetic on the device (migration.cu:90-117):
 * crop, image/illumination, minus z second difference; nz*nx*ny out. */
int fwi_export_image(fwi_ctx* ctx, float* out_interior);

/* Inversion::compute_gradient post-processing on the device
 * (inversion.cu:142-169): crop, gradient/illumination, masked minus z second
 * difference below zmask [m], scaled by 1/max|g|; *gmax receives max|g|. */
int fwi_export_gradient(fwi_ctx* ctx, float zmask, float* out_interior, float* gmax);

/* Residual r = obs - cal of the last recorded seismogram, kept on the device
 * as the backward data, and sum r^2 reduced in FP64 on the device
 * (Inversion::set_adjoint_source, inversion.cu:190-201; the reference sums
 * in sequential FP32 on the host). */
int fwi_residual(fwi_ctx* ctx, const float* obs, double* sum_sq);

/* ---- Batched shots -------------------------------------------------------
 * A batch context holds `nshots` shot slots that advance together: one step
 * kernel per time step updates every slot (shots of uneven receiver aperture
 * included), and the backward pass sums all slots into one image /
 * illumination / gradient.  The reference runs shots one after the other
 * (modeling_main.cpp:13-22, migration_main.cpp:13-19,
 * inversion.cu:123-140); a batch replaces that loop.  Models set with
 * fwi_set_model / fwi_set_model_interior go to every slot; the calls below
 * address one slot.  Entry points without a slot argument address slot 0. */
int fwi_create_batch(const fwi_grid* grid, int nshots, int device, fwi_ctx** out);
int fwi_shot_random_boundary(fwi_ctx* ctx, int slot, uint64_t seed, int shot, float amp, float vclamp);
int fwi_shot_random_boundary_blobs(fwi_ctx* ctx, int slot, uint64_t seed, int shot, float ratio, float var_vp,
                                   float vclamp);
int fwi_shot_set_source(fwi_ctx* ctx, int slot, int sIdx, int sIdy, int sIdz);
int fwi_shot_set_receivers(fwi_ctx* ctx, int slot, int nrec, const int* rIdx, const int* rIdy, const int* rIdz);
/* A slot that carries no shot: no source and no receivers, so its fields
 * stay zero and it adds nothing to the image, illumination or gradient of a
 * backward pass (the padding of a last, partial batch of shots; the
 * reference runs shot by shot and has no such slot).  fwi_shot_set_source
 * enables the slot again. */
int fwi_shot_disable(fwi_ctx* ctx, int slot);
int fwi_shot_get_seismogram(fwi_ctx* ctx, int slot, float* out);
int fwi_shot_set_backward_data(fwi_ctx* ctx, int slot, const float* data);
int fwi_shot_residual(fwi_ctx* ctx, int slot, const float* obs, double* sum_sq);
/* The slot's backward data (after fwi_shot_residual: the residual), nt*nrec
 * floats, t + r*nt -- the seismogram the reference copies back to the host
 * in set_adjoint_source (inversion.cu:192-200). */
int fwi_shot_get_backward_data(fwi_ctx* ctx, int slot, float* out);
int fwi_shot_get_volume(fwi_ctx* ctx, int slot, int which, float* out_padded);

/* Shots advanced by one block (1, 2 or 4; must divide nshots): the shots of
 * a block share the image / illumination read-modify-write. */
int fwi_set_shots_per_block(fwi_ctx* ctx, int kind, int sb);
int fwi_get_shots_per_block(fwi_ctx* ctx, int kind, int* sb);

/* Shots advanced by one thread: 2 = each thread advances one cell of two
 * consecutive slots (a shot pair; nshots even), on 64 x by tiles (by 4, 8 or
 * 16) and `zchunks` blocks along z.  The pair shares one image /
 * illumination read-modify-write (summed in the order of two shots per
 * block) and, in the interior, the velocity read.  1 = one shot per thread
 * (the shots-per-block setting applies).  Reference arithmetic only; a
 * backward pass with corrections runs pairs as two wave groups per block.
 * Batched counterpart of the per-shot launches of modeling.cu:314-321 and
 * migration.cu:72-80, which the reference runs one shot at a time. */
int fwi_set_shots_per_thread(fwi_ctx* ctx, int kind, int spt, int by, int zchunks);
int fwi_get_shots_per_thread(fwi_ctx* ctx, int kind, int* spt);

/* Autotune tile shape, z-chunking, prefetch depth and shots per block or
 * per thread for a pass kind (0 forward RBC, 1 forward ABC, 2 RTM backward,
 * 3 FWI backward) by timing `steps` steps of each candidate (backward
 * kinds: rounded up to a multiple of six, the deferred-imaging pattern of
 * whole passes), then the four fastest again with four times the steps;
 * cache_path (may be NULL) persists the choice.  The trial steps run on the
 * context's field buffers filled with noise (the kernels draw more power,
 * and rank differently, on non-zero fields); every field, image,
 * illumination and gradient buffer is zeroed afterwards: autotune before
 * setting up a pass.  Choosing several shots per block or per thread for a
 * backward kind changes the order in which the shots' image terms are
 * summed (equal within rounding). */
int fwi_autotune(fwi_ctx* ctx, int kind, int steps, const char* cache_path);

/* Current launch configuration: tile rows per block, rows per thread,
 * z-chunks, shots per block (1 for single-shot contexts). */
int fwi_get_tuning(fwi_ctx* ctx, int kind, int* by, int* rpt, int* zchunks);
int fwi_set_tuning(fwi_ctx* ctx, int kind, int by, int rpt, int zchunks);

/* Planes of loads in flight in the step kernel (1 or 2; part of the tuning,
 * chosen by fwi_autotune): with 2 the loads of plane z+2 are issued while
 * plane z is computed.  Two are available for one shot per block and the
 * tile shapes whose registers allow it; a backward pass with corrections
 * runs the one-plane instance of the same shape.  No reference counterpart. */
int fwi_set_prefetch_depth(fwi_ctx* ctx, int kind, int pd);
int fwi_get_prefetch_depth(fwi_ctx* ctx, int kind, int* pd);

/* Two time steps per launch (temporal blocking) for the random-boundary
 * forward pass (kind 0) and the RTM backward pass (kind 2); replaces pairs of
 * the per-step launches of modeling.cu:314-321 (compute_pressure) and
 * migration.cu:72-80 (compute_pressure + cross_correlation), with the same
 * results bit for bit.  A block of by x 64 threads (rpt rows each) computes
 * the first step on a 64 x (by*rpt) region and the second on its inner
 * 56 x (by*rpt - 8) cells; tiles: 16 x 2 and 12 x 3.  The forward pass
 * records the seismogram of both steps, the RTM pass injects the data of the
 * second step and adds both steps' image and illumination terms, all inside
 * the launch.  Used where available: one shot per block, volumes below 2 GiB,
 * each slot's receivers on at most four planes, and (RTM) no correction
 * flags; elsewhere, and for the last step of an odd step range, the one-step
 * launches run.  fwi_autotune chooses it when it is faster.
 * fwi_get_two_step: active = 1 when the next pass of that kind uses it,
 * -1 when requested but not available, 0 when off. */
int fwi_set_two_step(fwi_ctx* ctx, int kind, int enable, int by, int rpt, int zchunks);
int fwi_get_two_step(fwi_ctx* ctx, int kind, int* active, int* by, int* rpt, int* zchunks);

/* Capture the per-step launches of a full pass in a hipGraph (1) or launch
 * them on the stream (0, default).  A pass's graph is instantiated once and
 * replayed by later passes of the same kind and step range, until a setter
 * changes what its launches captured (wavelet, tuning, physics, imaging
 * cadence, arithmetic, geometry); fwi_graph_instantiations counts the
 * instantiations of the context (a diagnostic). */
int fwi_set_graph(fwi_ctx* ctx, int enable);
int fwi_graph_instantiations(fwi_ctx* ctx, int* count);

/* Device time of the last full pass's step kernels [ms], from HIP events
 * recorded on the context's stream, and the number of step launches. */
int fwi_last_pass_timing(fwi_ctx* ctx, double* ms, int* launches);

/* Record a HIP event pair around every step kernel of the following passes
 * (stream mode only) so that fwi_last_kernel_timing reports the summed
 * device time of the step kernels alone, without injection kernels. */
int fwi_set_kernel_timing(fwi_ctx* ctx, int enable);
int fwi_last_kernel_timing(fwi_ctx* ctx, double* ms, int* launches);

/* ---- Learned preconditioner ------------------------------------------------
 * No counterpart in the reference.  A 3-level 3-D U-Net (16/32/64 channels,
 * 3x3x3 convolutions, BF16 weights and activations, FP32 accumulation on the
 * matrix cores) maps (gradient, illumination, model) to a model update.
 * Inputs x3 are planar FP32 [3][B][D][H][W] (D, H, W divisible by 4; x
 * fastest); outputs / targets are FP32 [B][D][H][W]. */
typedef struct fwi_unet fwi_unet;
int fwi_unet_create(int device, uint64_t seed, fwi_unet** out);
int fwi_unet_destroy(fwi_unet* net);
const char* fwi_unet_last_error(const fwi_unet* net);
long fwi_unet_num_params(const fwi_unet* net);
int fwi_unet_get_params(fwi_unet* net, float* out);
int fwi_unet_set_params(fwi_unet* net, const float* in);
int fwi_unet_forward(fwi_unet* net, const float* x3, int B, int D, int H, int W, float* y);
/* MSE loss mean((y - target)^2) and the gradient of every parameter.
 * Deterministic: the gradient sums run in a fixed order (no float atomics),
 * so equal inputs and parameters give equal bits. */
int fwi_unet_gradients(fwi_unet* net, const float* x3, const float* target, int B, int D, int H, int W,
                       double* loss, float* grad);
/* One Adam step (beta 0.9 / 0.999) on the MSE loss. */
int fwi_unet_train_step(fwi_unet* net, const float* x3, const float* target, int B, int D, int H, int W, float lr,
                        double* loss);
/* One convolution layer alone (FP32 output, no bias/activation), for tests;
 * in is channels-last FP32 [B][D][H][W][Cin] rounded to BF16 on upload. */
int fwi_unet_conv_probe(fwi_unet* net, int layer, const float* in, int B, int D, int H, int W, float* out);
/* The backward kernels of one convolution layer alone, for tests: data
 * gradient dx = W^T dout ([B][D][H][W][Cin]) and weight gradient
 * dw = sum dout x ([Cout][27][Cin]) of the channels-last input `in`
 * (rounded to BF16 on upload; skip-concatenated layers take their two inputs
 * as the first and last channels) and the FP32 output gradient dout
 * ([B][D][H][W][Cout]). */
int fwi_unet_conv_grad_probe(fwi_unet* net, int layer, const float* in, const float* dout, int B, int D, int H, int W,
                             float* dx, float* dw);

/* Device-to-device copy bandwidth [GB/s, read + write] of a plain streaming
 * copy kernel over `bytes`, `reps` times: the measured reference the stencil
 * kernels' bandwidth is compared with.  No counterpart in the reference. */
int fwi_copy_bandwidth(fwi_ctx* ctx, long bytes, int reps, double* gbps);

/* Synchronise the context's stream. */
int fwi_sync(fwi_ctx* ctx);

/* ---- Allocation-free variant ------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight loops, hosts that forbid
 * hipMalloc): the residual of fwi_residual on caller device buffers, with no
 * context and no allocation.  obs and cal hold n floats in device memory;
 * cal is overwritten with r = obs - cal and *sum_sq (a device double)
 * receives sum r^2 reduced in FP64, with the same kernels, block count and
 * summation order as fwi_residual (equal inputs in the same order give equal
 * bits).  Runs asynchronously on `stream` (a hipStream_t; NULL = the null
 * stream) of the current device.  The per-block partial sums use a fixed
 * 8 KiB device array of the library, one per device, kept until the process
 * ends: calls that overlap on one device race.  n must be in
 * [1, FWI_NOALLOC_MAX_SAMPLES]; fwi_last_error(NULL) explains a failure. */
#define FWI_NOALLOC_MAX_SAMPLES (1L << 34)
int fwi_residual_noalloc(const float* obs, float* cal, long n, double* sum_sq, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FWI3D_H */
