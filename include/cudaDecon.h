/*
 * libcudaDecon.so - C ABI of the MI355X (gfx950) build.
 *
 * The first block reproduces, signature for signature, the exported API of
 * the reference library (src/linearDecon.h:243-317, implemented at
 * src/RL-Biggs-Andrews.cpp:453-855), so ctypes/cgo/JNI callers bind the
 * same symbols.  C++ default arguments of the reference header are kept for
 * C++ callers.  Errors are reported as the reference's return values (1 ok,
 * 0 failure); no function calls exit() or lets an exception cross the ABI.
 * The message of the last failure is available from cudaDecon_last_error().
 *
 * Not thread-safe, not re-entrant: one process-global session, as in the
 * reference (src/RL-Biggs-Andrews.cpp:459-466).  The current HIP device at
 * RL_interface_init time is used (the reference hard-codes device 0).
 *
 * The second block adds symbols with no counterpart in the reference.
 */
#ifndef CUDADECON_AMD_H
#define CUDADECON_AMD_H

#include <stddef.h>

#ifdef __cplusplus
#define CDN_DEFAULT(v) = v
extern "C" {
#else
#include <stdbool.h>
#define CDN_DEFAULT(v)
#endif

/* ---- reference API ---------------------------------------------------- */

/* replaces RL_interface_init (src/linearDecon.h:260-264;
 * src/RL-Biggs-Andrews.cpp:531-655).  Returns 1, or 0 if the OTF file cannot
 * be read or an FFT length is not supported. */
int RL_interface_init(int nx, int ny, int nz, float dr, float dz, float dr_psf, float dz_psf, float deskewAngle,
                      float rotationAngle, int outputWidth, bool bSkewedDecon, char* OTF_file_name);

/* replaces RL_interface (src/linearDecon.h:278-283;
 * src/RL-Biggs-Andrews.cpp:657-705).  `result` holds
 * get_output_nx()*get_output_ny()*get_output_nz() floats. */
int RL_interface(const unsigned short* const raw_data, int nx, int ny, int nz, float* result,
                 float* raw_deskewed_result, float background, bool bDoRescale, bool bSaveDeskewedRaw, int nIters,
                 int extraShift, int napodize CDN_DEFAULT(0), int nZblend CDN_DEFAULT(0),
                 float padVal CDN_DEFAULT(0), bool bDupRevStack CDN_DEFAULT(false),
                 bool bSkewedDecon CDN_DEFAULT(false));

/* replaces Deskew_interface (src/linearDecon.h:285-288;
 * src/RL-Biggs-Andrews.cpp:724-773). */
int Deskew_interface(const float* const raw_data, int nx, int ny, int nz, float dz, float dr, float deskewAngle,
                     float* const result, int outputWidth, int extraShift, float padVal CDN_DEFAULT(0));

/* replaces Affine_interface (src/linearDecon.h:290-292;
 * src/RL-Biggs-Andrews.cpp:776-814).  affMat: 16 floats, row-major 4x4. */
int Affine_interface(const float* const raw_data, int nx, int ny, int nz, float* const result, const float* affMat);

/* replaces Affine_interface_RA (src/linearDecon.h:294-296;
 * src/RL-Biggs-Andrews.cpp:816-855). */
int Affine_interface_RA(const float* const raw_data, int nx, int ny, int nz, float dx, float dy, float dz,
                        float* const result, const float* affMat);

/* declared by the reference (src/linearDecon.h:298-303; kernel
 * src/camcor.cu:80-108, intended context src/llspy.cpp:70-91) but never
 * exported by it.  init keeps the 3 parameter planes (a, b, offset; each
 * nx*ny floats, x fastest); camcor_interface corrects a uint16 stack:
 * out = v - offset - 0.9 a (1 - exp(-b (v_prev - offset))), v_prev the
 * previous plane (plane nz-1 for z = 0), truncated and saturated to uint16. */
int camcor_interface_init(int nx, int ny, int nz, const float* const camparam);
int camcor_interface(const unsigned short* const raw_data, int nx, int ny, int nz, unsigned short* const result);

/* replaces RL_cleanup (src/linearDecon.h:306; src/RL-Biggs-Andrews.cpp:707-715) */
void RL_cleanup(void);
/* replaces cuda_reset (src/linearDecon.h:307; src/RL-Biggs-Andrews.cpp:717-720) */
void cuda_reset(void);

/* replace get_output_nx/ny/nz (src/linearDecon.h:311-313;
 * src/RL-Biggs-Andrews.cpp:474-503) */
unsigned get_output_nx(void);
unsigned get_output_ny(void);
unsigned get_output_nz(void);

/* replaces get_version (src/linearDecon.h:316) */
const char* get_version(void);

/* ---- extensions (no counterpart in the reference) --------------------- */

/* Message of the last failed call ("" if none). */
const char* cudaDecon_last_error(void);

/* 1 if an FFT length is compiled into this library. */
int cudaDecon_fft_supported(int n);

/* findOptimalDimension (src/RL-Biggs-Andrews.cpp:33-43); n <= 1 returned as is. */
unsigned cudaDecon_find_optimal_dimension(unsigned n, int step);

/* Flat start of the later RL_interface calls (until RL_cleanup or a call
 * with enable = 0): the first prediction Y_0 is a volume filled with
 * `value` instead of the preprocessed stack, as the reference CLI's
 * --FlatStart does with the stack's median (src/linearDecon.cpp:1143-1145,
 * src/RL-Biggs-Andrews.cpp:305-316).  0 for a non-finite value. */
int cudaDecon_set_flat_start(int enable, float value);

/* Width of the result Deskew_interface writes for these arguments (its
 * sizing rule, src/RL-Biggs-Andrews.cpp:736-744, including the session's
 * output nz when outputWidth is 0), so callers can allocate nz*ny*width. */
unsigned cudaDecon_deskew_width(int nx, float dz, float dr, float deskewAngle, int outputWidth);

/* Session sizes: which = 0 output nz (before rotation), 1 deskewed nx,
 * 2 cropped nx, 3 FFT nx of the iteration.  0 without a session. */
unsigned cudaDecon_get_session_dims(int which);

/* OTF k-scales (kx, ky, kz) of the current session's iteration grid, as
 * RL_interface_init computes them (src/RL-Biggs-Andrews.cpp:634-642). */
int cudaDecon_get_session_kscale(float* ks3);

/* Unnormalised 3-D R2C / C2R of host arrays: real (nz, ny, nx) <->
 * complex64 interleaved (nz, ny, nx/2+1).  C2R ignores the imaginary parts
 * of the kx = 0 and kx = nx/2 bins. */
int cudaDecon_fft_r2c(const float* in, int nx, int ny, int nz, float* out);
int cudaDecon_fft_c2r(const float* in, int nx, int ny, int nz, float* out);

/* One filterGPU call (src/RLgpuImpl.cu:156-225) on a host volume, in place:
 * v = C2R(H . R2C(v / N)), H (or conj H) interpolated from an OTF image laid
 * out as its TIFF file (width, height, depth floats) with k-scales kx, ky, kz. */
int cudaDecon_filter(float* vol, int nx, int ny, int nz, const float* otf, int otf_width, int otf_height,
                     int otf_depth, float kxscale, float kyscale, float kzscale, int conj);

/* One-step Wiener filter, the CLI's alternative to RL (wienerfilter,
 * src/linearDecon.cpp:116-168, called at :1116-1128), on the GPU, in place:
 * v = C2R(conj(H) R2C(v - background) / (|H|^2 + wiener^2)
 *         * (1 - kr / rcutoff)) / N  for kr = |(kx dkx, ky dky)| <= rcutoff,
 * 0 beyond.  H: rotationally averaged OTF image (width 2 nzotf, height nr),
 * k-scales as for RL; dkx = 1 / (dr nx), dky = 1 / (dr ny); the CLI uses
 * rcutoff = 2 NA / wavelength. */
int cudaDecon_wiener(float* vol, int nx, int ny, int nz, const float* otf, int otf_width, int otf_height,
                     float kxscale, float kyscale, float kzscale, float dkx, float dky, float rcutoff,
                     float wiener, float background);

/* Autotuner of the FFT launch shapes.  For every length used by the grids
 * (dims_zyx: 3 ints nz, ny, nx per grid) it times candidate shapes of the x
 * passes (workgroup size) and of the y/z column passes (columns per tile,
 * cached or non-temporal loads) on the largest grid using the length,
 * compiles them at run time, and installs the fastest for all later calls
 * of this process (results are bitwise unchanged).  use_cache: reuse and
 * extend the decisions kept under the cache directory.  report (nullable):
 * a JSON summary of the timings.  Do not run while other work of this
 * library is in flight. */
int cudaDecon_autotune(const int* dims_zyx, int count, int reps, int use_cache, char* report, int report_len);
/* back to the built-in shapes for every length */
int cudaDecon_autotune_clear(void);
/* 1 if length n has a tuned shape; its first nfields parameters (0 / -1 =
 * built-in) go to fields: columns per column tile, threads per x-pass
 * workgroup, non-temporal column loads, workgroup-to-XCD renumbering of the
 * column passes (0 none, 1 tiles narrower than a cache line, 2 every tile),
 * threads per column line, the z filter's minimum waves per SIMD, OTF table
 * rows the z filter stages in LDS per tile (-1 none). */
int cudaDecon_fft_variant(int n, int* fields, int nfields);

/* RL-Biggs-Andrews engine on HBM-resident state (the iteration of
 * RL_interface without pre/post-processing).  The OTF is given as its
 * image (see cudaDecon_filter) and the k-scales of the grid. */
typedef struct cudaDecon_rl cudaDecon_rl;
cudaDecon_rl* cudaDecon_rl_create(int nx, int ny, int nz, const float* otf, int otf_width, int otf_height,
                                  int otf_depth, float kxscale, float kyscale, float kzscale);
/* One engine over nstacks stacks of one shape (tiles, patches, time
 * points): stored back to back, every pass of an iteration is one launch
 * over all of them, each stack keeping its own acceleration factor; each
 * stack's result equals a one-stack engine's.  ny nz must be a multiple of
 * the x passes' rows per workgroup (the call fails naming it otherwise).
 * No counterpart in the reference (its library holds one stack). */
cudaDecon_rl* cudaDecon_rl_create_stacks(int nx, int ny, int nz, int nstacks, const float* otf, int otf_width,
                                         int otf_height, int otf_depth, float kxscale, float kyscale,
                                         float kzscale);
int cudaDecon_rl_stacks(cudaDecon_rl* h);
/* stage / read stack s alone (host memory, nx ny nz floats) */
int cudaDecon_rl_set_raw_stack(cudaDecon_rl* h, int s, const float* raw);
int cudaDecon_rl_get_stack(cudaDecon_rl* h, int s, float* X);
/* preprocessed stack (also X_0), host memory; all stacks, one after another */
int cudaDecon_rl_set_raw(cudaDecon_rl* h, const float* raw);
/* seed for Y_0 (the FlatStart hook, src/RL-Biggs-Andrews.cpp:305-316) */
int cudaDecon_rl_set_seed(cudaDecon_rl* h, const float* seed);
/* iterations k0 .. k0+n-1 of every stack; lambdas (n x nstacks, stack
 * fastest, nullable) receives the factor used at each k; elapsed_ms
 * (nullable) the device time of the n iterations. */
int cudaDecon_rl_iterate(cudaDecon_rl* h, int n, int k0, int seeded, double* lambdas, double* elapsed_ms);
/* X of all stacks, one after another */
int cudaDecon_rl_get(cudaDecon_rl* h, float* X);
/* per-pass device time: enable != 0 turns event timing on (and clears the
 * sums); pass_ms / pass_calls (7 entries: x-R2C, y-fwd, z-filter, y-inv,
 * x-LRcore, x-update, lambda) receive the sums so far (nullable). */
int cudaDecon_rl_timing(cudaDecon_rl* h, int enable, double* pass_ms, long* pass_calls);
/* bytes of one half-spectrum row pitch (complex elements) and volume sizes
 * (of one stack) */
int cudaDecon_rl_geometry(cudaDecon_rl* h, long* nvox, long* nspec_alg, long* pitch);
/* Stage a sparse stack without a full-volume transfer: every voxel is set
 * to `background`, then the host block (bz, by, bx) is placed at offset
 * (oz, oy, ox); X_0 := the stack.  get_block reads a block of X back.
 * These, stats, the pipeline below (set_raw_u16, deskew, affine) and
 * snapshot / distance are defined for one-stack engines and fail on others
 * (warm_start seeds every stack). */
int cudaDecon_rl_set_raw_block(cudaDecon_rl* h, const float* block, int bx, int by, int bz, int ox, int oy, int oz,
                               float background);
int cudaDecon_rl_get_block(cudaDecon_rl* h, float* out, int bx, int by, int bz, int ox, int oy, int oz);
/* FP64 sum, min, max of the staged stack (out6[0..2]) and of X (out6[3..5]). */
int cudaDecon_rl_stats(cudaDecon_rl* h, double* out6);
/* Plain 16-byte streaming copy of the stack into a scratch volume, n times;
 * *ms receives the average device time of one copy (bandwidth reference). */
int cudaDecon_rl_copy_probe(cudaDecon_rl* h, int n, double* ms);
void cudaDecon_rl_destroy(cudaDecon_rl* h);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid allocation): the no-allocation variant of cudaDecon_rl_create +
 * cudaDecon_rl_set_raw + cudaDecon_rl_iterate(n_iters from k = 0) +
 * cudaDecon_rl_get, with the same results.  d_raw: the preprocessed stack
 * (device, nz*ny*nx floats, x fastest, read only); d_out: device result of
 * the same size; otf and k-scales as cudaDecon_rl_create.  Every other
 * device buffer is carved from a static array of 256 MiB that each GPU the
 * library runs on holds until the process ends (about 20 bytes per voxel:
 * stacks up to about 13 M voxels, e.g. 64 x 256 x 768); larger stacks and
 * FFT lengths that are not built in are refused with an error.  Runs on
 * `stream` (NULL: the null stream) and returns after it finished.  Calls on
 * one GPU share the array and are serialised by the library. */
int cudaDecon_rl_run_noalloc(const float* d_raw, float* d_out, int nx, int ny, int nz, const float* otf,
                             int otf_width, int otf_height, int otf_depth, float kxscale, float kyscale,
                             float kzscale, int n_iters, void* stream);

/* Device memory for the entry points that take device pointers: allocate
 * (NULL on failure), free, and copy (kind 1 host to device, 2 device to host,
 * 3 device to device; synchronous). */
void* cudaDecon_device_malloc(size_t bytes);
int cudaDecon_device_free(void* p);
int cudaDecon_device_copy(void* dst, const void* src, size_t bytes, int kind);

/* Batch of independent stacks of uneven shape (time points, channels, tiles).
 * dims_zyx: 3 ints (nz, ny, nx) per stack; kscales: 3 floats per stack (the
 * k-scales RL_interface_init would compute for that stack's grid).
 * run: `iters` iterations of every stack from its staged stack:
 *   mode 0 = each stack's own engine, one stack at a time on one stream;
 *   1 = stacks round-robin over nstreams streams; 2 = the mode-1 schedule
 *   captured once into a hipGraph and replayed;
 *   3 = grouped: every pass is one launch per distinct transform length,
 *   spanning all stacks of that length whatever their other sizes (about
 *   3 nx + 4 ny + 2 nz + 2 launches per iteration for the distinct lengths);
 *   4 = the mode-3 schedule captured once into a hipGraph and replayed.
 * Every mode gives each stack bitwise the result of a one-stack engine.
 * *ms receives the device time of the whole batch.  set_raw / set_raw_block
 * reject stacks holding NaN or Inf. */
typedef struct cudaDecon_batch cudaDecon_batch;
cudaDecon_batch* cudaDecon_batch_create(int n, const int* dims_zyx, const float* otf, int otf_width, int otf_height,
                                        int otf_depth, const float* kscales, int nstreams);
/* As cudaDecon_batch_create with several OTFs (channels): ntables OTF images
 * laid out as their TIFF files, back to back in otf_data, otf_whd holding
 * (width, height, depth) of each; table_of[i] is stack i's table. */
cudaDecon_batch* cudaDecon_batch_create_tables(int n, const int* dims_zyx, int ntables, const float* otf_data,
                                               const int* otf_whd, const int* table_of, const float* kscales,
                                               int nstreams);
/* Acceleration factors of stack i at iterations 0 .. iters-1 of the last run
 * (0 for k < 2). */
int cudaDecon_batch_lambdas(cudaDecon_batch* h, int i, int iters, double* out);
/* Grouped modes: column passes on a persistent grid of at most
 * max_workgroups workgroups that walk the tiles (0: one workgroup per tile). */
int cudaDecon_batch_set_grid_cap(cudaDecon_batch* h, long max_workgroups);
/* Grouped modes: cut the stacks (ordered by x length) into n ranges of about
 * equal voxel count, each grouped on its own and run on its own stream, so
 * that passes of different ranges overlap (1: one schedule on one stream;
 * at most the batch's nstreams). */
int cudaDecon_batch_set_partitions(cudaDecon_batch* h, int n);
/* Kernel launches per iteration of the grouped modes. */
int cudaDecon_batch_launches(cudaDecon_batch* h);
int cudaDecon_batch_set_raw(cudaDecon_batch* h, int i, const float* raw);
int cudaDecon_batch_set_raw_block(cudaDecon_batch* h, int i, const float* block, int bx, int by, int bz, int ox,
                                  int oy, int oz, float background);
int cudaDecon_batch_run(cudaDecon_batch* h, int iters, int mode, double* ms);
int cudaDecon_batch_get(cudaDecon_batch* h, int i, float* X);
void cudaDecon_batch_destroy(cudaDecon_batch* h);

/* Device-resident pipeline around the engine (no host round trip):
 * set_raw_u16 stages a uint16 stack with camera-offset subtraction
 * (max(v - offset, 0), the background kernel src/RLgpuImpl.cu:124-132);
 * deskew shears the current X (deskew_kernel, src/geometryTransform.cu:4-48)
 * with the factor of RL_interface_init (width: outputWidth, or
 * findOptimalDimension(nx + floor(nz dz |cos| / dr)) when 0); affine
 * resamples the last result (the deskewed volume, else X) with a row-major
 * 4x4 matrix (affine_GPU, src/geometryTransform.cu:183-290).  post_* read
 * the last deskew / affine result. */
int cudaDecon_rl_set_raw_u16(cudaDecon_rl* h, const unsigned short* raw, float offset);
int cudaDecon_rl_deskew(cudaDecon_rl* h, float dz, float dr, float deskewAngle, int outputWidth, int extraShift,
                        float padVal, int* nx_out);
int cudaDecon_rl_affine(cudaDecon_rl* h, const float* affMat);
int cudaDecon_rl_post_dims(cudaDecon_rl* h, int* nx, int* ny, int* nz);
int cudaDecon_rl_post_get(cudaDecon_rl* h, float* out);
/* FP64 sum, min, max of the last deskew / affine result */
int cudaDecon_rl_post_stats(cudaDecon_rl* h, double* out3);
/* A block of the deskew (buf 0) or affine (buf 1, after cudaDecon_rl_affine
 * of a deskew result) result, bx*by*bz floats from (ox, oy, oz). */
int cudaDecon_rl_post_get_block(cudaDecon_rl* h, int buf, float* out, int bx, int by, int bz, int ox, int oy,
                                int oz);
/* FP64 sum, min, max of an engine buffer: 0 the staged stack, 1 X, 2 Y (the
 * seed after cudaDecon_rl_set_seed or cudaDecon_rl_warm_start), 3 G. */
int cudaDecon_rl_buffer_stats(cudaDecon_rl* h, int which, double* out3);

/* Reference copy of the engine's current X (snapshot) and the relative L2
 * distance ||V - ref|| / ||ref|| (FP64 sums) of V = the current X
 * (which = 0) or the Y_0 seed buffer (which = 1). */
int cudaDecon_rl_snapshot(cudaDecon_rl* h);
int cudaDecon_rl_distance(cudaDecon_rl* h, int which, double* rel);

/* Learned warm start: a 3-level 3-D U-Net (16/32/64 channels, 3x3x3
 * convolutions, ReLU, average pooling, nearest upsampling with additive
 * skips, 1x1x1 head plus a skip add of the input) in BF16 on the matrix
 * cores with FP32 accumulation.  Input u = stack / mean(stack) per sample;
 * output y (normalised); RL seed = max(y, 0) * mean(stack).  Volumes are
 * (nb, nz, ny, nx) float arrays with nz, ny, nx multiples of 4.
 * Parameters are one FP32 vector: per 3x3x3 layer W[27][cin][cout] then
 * b[cout] (layers 1->16, 16->32, 32->64, 64->32, 32->16), then the head
 * Wh[16], bh. */
typedef struct cudaDecon_net cudaDecon_net;
/* He-normal weights from `seed`, zero biases and head (output = input). */
cudaDecon_net* cudaDecon_net_create(unsigned long long seed);
long cudaDecon_net_num_params(void);
int cudaDecon_net_get_params(cudaDecon_net* h, float* params);
/* also clears the Adam state */
int cudaDecon_net_set_params(cudaDecon_net* h, const float* params);
/* mode 0: y, mode 1: the RL seed; out has the shape of raw.  Modes 2 / 3:
 * the same through the inference path (full-size e1 recomputed in the head
 * instead of stored; what cudaDecon_rl_warm_start runs). */
int cudaDecon_net_forward(cudaDecon_net* h, const float* raw, int nb, int nz, int ny, int nx, int mode, float* out);
/* Copy an internal tensor of the last forward / loss_grad call to host as
 * float (channels-last).  which: 0 e1, 1 pool(e1), 2 e2, 3 pool(e2), 4 e3,
 * 5 d3, 6 up(d3)+e2, 7 d2, 8 dL/dy, 9.. pre-activation gradients of the
 * layers and the gradients flowing between them (see the implementation).
 * Returns the element count (out may be null to ask), -1 on error. */
long cudaDecon_net_fetch(cudaDecon_net* h, int which, float* out);
/* loss = mean((y - target / mean(raw))^2) and its gradient w.r.t. every
 * parameter (num_params floats) */
int cudaDecon_net_loss_grad(cudaDecon_net* h, const float* raw, const float* target, int nb, int nz, int ny, int nx,
                            double* loss, float* grad);
/* one Adam step (lr, betas 0.9 / 0.999, eps 1e-8) on a host batch */
int cudaDecon_net_step(cudaDecon_net* h, const float* raw, const float* target, int nb, int nz, int ny, int nx,
                       float lr, double* loss);
/* append `count` (stack, target) pairs of one shape to the device training set */
int cudaDecon_net_add_samples(cudaDecon_net* h, const float* raw, const float* target, int count, int nz, int ny,
                              int nx);
/* `steps` Adam steps on random batches of the training set; losses (steps
 * floats, nullable); elapsed_ms (nullable) the device time */
int cudaDecon_net_train(cudaDecon_net* h, int steps, int batch, float lr, unsigned long long seed, float* losses,
                        double* elapsed_ms);
void cudaDecon_net_destroy(cudaDecon_net* h);
/* Seed Y_0 of an RL engine with the network's output on the engine's
 * staged stack (all on the device); then iterate with seeded = 1.
 * elapsed_ms (nullable): device time of the network. */
int cudaDecon_rl_warm_start(cudaDecon_rl* h, cudaDecon_net* net, double* elapsed_ms);

/* Synchronise the current device. */
int cudaDecon_device_sync(void);

#ifdef __cplusplus
}
#endif
#undef CDN_DEFAULT

#endif /* CUDADECON_AMD_H */
