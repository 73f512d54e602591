/*
 * lattibol_amd.h -- C ABI of the MI355X-native lattice-Boltzmann core.
 *
 * Two groups of entry points:
 *
 * 1. Drop-in replacements for the reference's extern "C" launchers, same
 *    names, same argument order and meaning, reference semantics (in-place
 *    swap streaming, full-way bounce-back, Zou-He inlet/outlet, D2Q9
 *    Shan-Chen).  The Data structs keep the reference's field order
 *    (lattibol.h:29-97); HIP's dim3 and hipStream_t are replaced by the
 *    layout-identical lb_dim3 and an opaque lb_stream_t so this header needs
 *    no GPU headers.  grid*block must equal the subdomain shape, as in the
 *    reference (simulation.cu:43-44); the library validates that and then
 *    launches its own shapes.  These functions return void like the
 *    reference; lb_last_error() reports the first failure since the last call.
 *
 * 2. New int-returning entry points (0 = success, >0 a hipError_t, <0 an
 *    LB_ERR_* code): the reference step loop as a host driver (lb_refsim_*),
 *    the fused single-sweep pull kernels with FP32/FP16/BF16/FP64 storage,
 *    halo planes, FP64 reductions, hipGraph-captured stepping (lb_*), the
 *    batched multi-lattice solver (lb_batch_*) and the U-Net surrogate
 *    (lb_unet_*).  Device memory is owned by opaque handles; every handle is
 *    single-threaded and stream-ordered on its own HIP stream.
 */
#ifndef LATTIBOL_AMD_H
#define LATTIBOL_AMD_H

#include <stddef.h>
#include <stdint.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Types shared with the reference                                         */
/* ---------------------------------------------------------------------- */
typedef struct { unsigned int x, y, z; } lb_dim3;   /* layout of HIP's dim3 */
typedef void *lb_stream_t;                            /* a hipStream_t       */

/* lattibol.h:29-63 */
typedef struct {
    float *ux, *uy, *rho;
    bool *solid, *solid_bound;
    float *f0, *f1, *f2, *f3, *f4;
    float *f5, *f6, *f7, *f8;
    float *border_f1, *border_f5, *border_f8;
    float *border_f3, *border_f6, *border_f7;
    float *nb_border_f1, *nb_border_f5, *nb_border_f8;
    float *nb_border_f3, *nb_border_f6, *nb_border_f7;
    int width, height;
    int deviceCount, dev;
    float *plotBuffer;
    float *valueCur;
    lb_stream_t stream;
    float *psi;
    float *ueq_x, *ueq_y;
} DataD2Q9;

/* lattibol.h:65-97 */
typedef struct {
    float *ux, *uy, *uz, *rho;
    bool *solid, *solid_bound;
    float *f0, *f1, *f2, *f3, *f4;
    float *f5, *f6, *f7, *f8, *f9;
    float *f10, *f11, *f12, *f13, *f14;
    float *f15, *f16, *f17, *f18;
    float *border_f1, *border_f7, *border_f8, *border_f9, *border_f10;
    float *border_f2, *border_f11, *border_f12, *border_f13, *border_f14;
    float *nb_border_f1, *nb_border_f7, *nb_border_f8, *nb_border_f9, *nb_border_f10;
    float *nb_border_f2, *nb_border_f11, *nb_border_f12, *nb_border_f13, *nb_border_f14;
    int width, height, depth;
    int deviceCount, dev;
    float *plotBuffer;
    float *valueCur;
    lb_stream_t stream;
} DataD3Q19;

/* ---------------------------------------------------------------------- */
/* Error codes                                                             */
/* ---------------------------------------------------------------------- */
enum {
    LB_OK = 0,
    LB_ERR_INVALID = -1,      /* bad argument / null pointer               */
    LB_ERR_SHAPE = -2,        /* grid*block does not match the lattice     */
    LB_ERR_UNSUPPORTED = -3,  /* configuration not implemented             */
    LB_ERR_NO_DEVICE = -4,    /* no usable GPU                             */
    LB_ERR_NOT_READY = -5,    /* state missing (e.g. macros never written)  */
    LB_ERR_SCRATCH = -6       /* work exceeds the fixed scratch of a
                                 non-allocating entry point                */
};
int lb_last_error(void);                 /* returns and clears the sticky error */
const char *lb_error_string(int code);
int lb_device_count(void);
/* Device-memory plumbing for hosts without a GPU framework. */
int lb_device_alloc(size_t bytes, void **out);
int lb_device_free(void *p);
int lb_memcpy_h2d(void *dst, const void *src, size_t bytes);
int lb_memcpy_d2h(void *dst, const void *src, size_t bytes);
int lb_device_synchronize(void);
/* A plain 16-byte-vector copy kernel of `bytes` (multiple of 16), launched
 * repeats times after one warm launch; average ms per copy.  It reads and
 * writes exactly `bytes`: the calibration point for HBM traffic counters. */
int lb_copy_device(void *dst, const void *src, size_t bytes, int repeats, float *ms_per_copy);

/* ---------------------------------------------------------------------- */
/* 1. Reference entry points (reference semantics, legacy default stream)  */
/* ---------------------------------------------------------------------- */
/* streaming.h:31-35 */
void streaming_d2q9(DataD2Q9 *data, lb_dim3 grid, lb_dim3 block);
void streaming_d3q19(DataD3Q19 *data, lb_dim3 grid, lb_dim3 block);
/* collision.h:27-36 (u_min..rho_max are accepted and unused, as in the
 * reference; the D2Q9 Shan-Chen second pass runs only for G < 0) */
void collision_d2q9(float *plot_data, DataD2Q9 *data, float w0, float w1, float w5,
                    float tau_inv, float ux_in, float uy_in, float rho_out, float u_min,
                    float u_max, float rho_min, float rho_max, lb_dim3 grid, lb_dim3 block,
                    float Fx, float Fy, float G, float Gads);
void collision_d3q19(float *plot_data, DataD3Q19 *data, float w0, float w1, float w7,
                     float tau_inv, float ux_in, float uy_in, float uz_in, float rho_out,
                     float u_min, float u_max, float rho_min, float rho_max, lb_dim3 grid,
                     lb_dim3 block);
/* get_border.h:29-33 */
void get_border_d2q9(DataD2Q9 *data, lb_dim3 grid, lb_dim3 block);
void get_border_d3q19(DataD3Q19 *data, lb_dim3 grid, lb_dim3 block);
/* copy_border.h:30-47.  The reference passes cudaStream_t& (a pointer at the
 * ABI level); the D3Q19 variant had C++ linkage there and is extern "C" here. */
void copy_border_devices_d2q9(float *nb_border_f1, float *nb_border_f5, float *nb_border_f8,
                              float *nb_border_f3, float *nb_border_f6, float *nb_border_f7,
                              float *border_f1, float *border_f5, float *border_f8,
                              float *border_f3, float *border_f6, float *border_f7,
                              int height, lb_stream_t *s);
void copy_border_devices_d3q19(float *nb_border_f1, float *nb_border_f7, float *nb_border_f8,
                               float *nb_border_f9, float *nb_border_f10, float *nb_border_f2,
                               float *nb_border_f11, float *nb_border_f12, float *nb_border_f13,
                               float *nb_border_f14, float *border_f1, float *border_f7,
                               float *border_f8, float *border_f9, float *border_f10,
                               float *border_f2, float *border_f11, float *border_f12,
                               float *border_f13, float *border_f14, int height, int depth,
                               lb_stream_t *s);
/* apply_border.h:25-29 */
void apply_border_d2q9(DataD2Q9 *data, lb_dim3 grid, lb_dim3 block);
void apply_border_d3q19(DataD3Q19 *data, lb_dim3 grid, lb_dim3 block);
/* read_value.h:27-31 (synchronises the device, like the reference) */
void read_value_2d(float *plot_data, DataD2Q9 *data, int x, int y, int z);
void read_value_3d(float *plot_data, DataD3Q19 *data, int x, int y, int z);

/* ---------------------------------------------------------------------- */
/* 1b. The reference host driver: Simulation{D2Q9,D3Q19} (simulation_d2q9.cu
 *     :52-560, simulation_d3q19.cu:50-436) as an opaque handle.           */
/* ---------------------------------------------------------------------- */
typedef struct {
    int deviceCount;              /* number of x-slab subdomains             */
    int width, height, depth;     /* full lattice; depth = 1 selects D2Q9    */
    float ux, uy, uz, rho;        /* inlet velocity, outlet / initial density */
    float viscosity;              /* tau_inv = 1/(3 nu + 0.5)                */
    float Fx, Fy, Fz;             /* body force (D2Q9)                       */
    float G, Gads;                /* Shan-Chen coupling / adhesion (D2Q9)    */
    int zou_he;                   /* 1 = reference inlet/outlet (its launchers always
                                     apply them); 0 = x periodic over the ring */
    int cursor_x, cursor_y, cursor_z;
} lb_ref_params;
typedef struct lb_refsim lb_refsim;
int lb_refsim_create(const lb_ref_params *p, const uint8_t *solids, lb_refsim **out);
int lb_refsim_destroy(lb_refsim *s);
/* runLBM(): updateSteps reference steps (stream, halo, collide, probe). */
int lb_refsim_run(lb_refsim *s, int updateSteps);
/* valueCur after the last step: (ux, uy, uz, rho) at the cursor. */
int lb_refsim_value_cur(lb_refsim *s, float out[4]);
/* Copies subdomain dev to host: f (Q*N), ux, uy, uz (NULL for D2Q9 ok), rho,
 * plot (8N floats for D2Q9, 4N for D3Q19).  Any pointer may be NULL. */
int lb_refsim_download(lb_refsim *s, int dev, float *f, float *ux, float *uy, float *uz,
                       float *rho, float *plot);
/* Access to the reference-layout structs (device pointers inside). */
int lb_refsim_data_d2q9(lb_refsim *s, int dev, DataD2Q9 *out);
int lb_refsim_data_d3q19(lb_refsim *s, int dev, DataD3Q19 *out);

/* ---------------------------------------------------------------------- */
/* 2. Fused pull-scheme solver                                              */
/* ---------------------------------------------------------------------- */
enum { LB_BC_PERIODIC = 0, LB_BC_WALL = 1, LB_BC_MOVING_WALL = 2, LB_BC_INLET_EQ = 3,
       LB_BC_OUTFLOW = 4 };
/* F32 / F64 store f_i and follow the reference expressions; F32S / F16 / BF16
 * store s_i = f_i - w_i and run the collision in deviation form (FP32). */
enum { LB_STORE_F32 = 0, LB_STORE_F16 = 1, LB_STORE_BF16 = 2, LB_STORE_F64 = 3, LB_STORE_F32S = 4 };
/* Device layout of the populations.  SOA: [i][z][y][x]; ROWS: [z][y][i][x]
 * (the Q populations of one x-row back to back).  The host-side population
 * API (lb_set_populations / lb_get_populations) always uses SOA order. */
enum { LB_LAYOUT_SOA = 0, LB_LAYOUT_ROWS = 1 };

/* Faces are 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z in storage index space (D2Q9 rows
 * grow downward as in the reference, physical y = -row).  Populations use
 * the reference numbering (streaming.cu, collision.cu). */
typedef struct {
    int lattice;            /* 9 or 19 */
    int nx, ny, nz;         /* nz = 1 for D2Q9 */
    int storage;            /* LB_STORE_* */
    float tau_inv;
    int bc[6];
    float bc_u[6][3];       /* physical wall / inlet velocity */
    float bc_rho[6];        /* inlet density */
    float force[3];         /* body force, velocity-shift forcing (collision.cu:281-284) */
    float sc_G;             /* Shan-Chen coupling (D3Q19), 0 = off */
    float sc_rho0;          /* psi = rho0 (1 - exp(-rho/rho0)) */
    float sc_psi_solid;
    int bb_full;            /* 1: full-way bounce-back at solid nodes (reference) */
    int zou_he;             /* 1: reference Zou-He inlet x=0 / outlet x=nx-1 */
    float zh_ux_in, zh_rho_out;
    int halo_z;             /* D3Q19 periodic z through packed halo planes:
                               pack (fused in the sweep) -> copy -> unpack (fused) */
    int block_x, block_y;   /* launch shape; 0 = autotune */
    int xcd_remap;          /* -1 autotune, 0 off, 1 on */
    int graph_chunk;        /* steps per captured hipGraph; 0 = default */
    int rest_closure;       /* 1: feq_0 = rho - sum_{i>0} feq_i (exact mass balance);
                               0: the reference's equilibrium for every direction */
    int mem_mode;           /* population loads/stores: -1 autotune, 0 plain,
                               1 non-temporal stores, 2 non-temporal loads + stores */
    int layout;             /* -1 autotune, LB_LAYOUT_SOA, LB_LAYOUT_ROWS */
    int round_mode;         /* FP16 / BF16 stores: 0 default = stochastic rounding for
                               Shan-Chen lattices (nearest-even there loses mass or
                               stalls the phase separation) and nearest-even otherwise;
                               1 stochastic rounding (unbiased; counter-based random
                               bits per step, node and population, reproducible);
                               2 nearest-even always */
    int node_words;         /* -1 autotune, 0 flag byte (+ link mask for nodes next to
                               walls / solids), 1 one packed 32-bit word per node */
    int time_block;         /* steps per launch through an LDS tile (temporal blocking,
                               for small launch-latency-bound lattices): -1 autotune,
                               0 off, 4 or 8.  D2Q9 FP32 only, faces periodic / wall /
                               moving wall, no force, Shan-Chen or Zou-He, nx and ny
                               >= 16 + time_block; else 0 (a fixed value that cannot
                               apply is LB_ERR_UNSUPPORTED) */
    int sc_sweep;           /* Shan-Chen D3Q19 step: -1 autotune, 0 density pass + sweep
                               (two launches, psi through HBM), 1 or 2 one z-marching
                               sweep with the psi planes in LDS (one launch), keeping
                               the pulled populations in registers (1) or pulling them
                               again from L2 for the sweep (2), 3 one z-marching sweep
                               over whole x-rows (row-slab tiles, no frame column) for
                               fully periodic lattices; all bit-identical.
                               1 / 2 / 3 need Shan-Chen D3Q19, FP32 / FP16 / BF16 storage
                               and plain accesses (mem_mode -1 or 0), else
                               LB_ERR_UNSUPPORTED.  3 also needs six periodic faces, no
                               halo planes, Zou-He faces or solid nodes (lb_set_solid
                               with a solid node then fails; under -1 the two passes
                               take over), nx <= 512 and 19 N sizeof(population) < 4 GiB;
                               block_x is then its workgroup (384, 512, 768 or 1,024 threads)
                               and block_y its rows per tile */
    int sc_zchunk;          /* z-planes per workgroup of the one-pass sweep, 0 = autotune */
} lb_config;

typedef struct lb_ctx lb_ctx;
int lb_create(const lb_config *cfg, lb_ctx **out);
int lb_destroy(lb_ctx *ctx);
int lb_get_config(lb_ctx *ctx, lb_config *out);     /* with autotuned fields */
/* Solid mask (host, N bytes, 1 = solid).  Also rebuilds the node flags. */
int lb_set_solid(lb_ctx *ctx, const uint8_t *solid);
/* Equilibrium from host fields (N floats each; uz may be NULL), computed on
 * the GPU; writes both population buffers and clears the step counter. */
int lb_init_equilibrium(lb_ctx *ctx, const float *rho, const float *ux, const float *uy,
                        const float *uz);
/* Same from device fields. */
int lb_init_equilibrium_device(lb_ctx *ctx, const float *d_rho, const float *d_ux,
                               const float *d_uy, const float *d_uz);
int lb_init_uniform(lb_ctx *ctx, float rho, float ux, float uy, float uz);
/* Raw populations in the storage format (Q*N elements, [i][z][y][x] order
 * whatever the device layout), host side. */
int lb_set_populations(lb_ctx *ctx, const void *host);
int lb_get_populations(lb_ctx *ctx, void *host);
/* nsteps time steps; the last one also writes rho,u (the macros).  The
 * steps run as captured hipGraphs: whole chunks of graph_chunk steps, then
 * one graph for the rest of the call (cached per length; the first call of a
 * new length captures and instantiates it). */
int lb_step(lb_ctx *ctx, int nsteps);
/* Captures and instantiates the graphs a following lb_step(ctx, nsteps) will
 * launch, without running anything (keeps capture out of a timed loop). */
int lb_step_prepare(lb_ctx *ctx, int nsteps);
int lb_sync(lb_ctx *ctx);
/* Macros of the last step to host (N floats each, any may be NULL). */
int lb_get_macros(lb_ctx *ctx, float *rho, float *ux, float *uy, float *uz);
/* Device pointer to the macro planes (rho, ux, uy, uz; 4*N floats). */
int lb_macros_device(lb_ctx *ctx, float **out);
/* FP64 totals over fluid nodes of the current populations:
 * out = {mass, momentum_x, momentum_y, momentum_z} (physical axes). */
int lb_reduce(lb_ctx *ctx, double out[4]);
/* ||u - u_snap|| / ||u|| over fluid nodes (FP64), then u_snap := u.  The
 * first call after init returns 1.0 (no snapshot yet). */
int lb_residual(lb_ctx *ctx, double *out);
/* (ux, uy, uz, rho) at one node from the last macros. */
int lb_read_value(lb_ctx *ctx, int x, int y, int z, float out[4]);
int lb_steps_done(lb_ctx *ctx, long long *out);
/* Timing helper: average duration (ms) of one step kernel, measured with HIP
 * events over n back-to-back launches on the handle's stream (no macros). */
int lb_time_step_kernel(lb_ctx *ctx, int n, float *ms_per_launch);
/* Halo buffers of the z-slab for the next step (device pointers): the send
 * planes that sweep packs (the five +z / five -z populations of the top /
 * bottom plane, 5*nx*ny elements each) and the recv planes it unpacks.  Two
 * buffers alternate step by step: the sweep packs into one and unpacks the
 * other, which the previous sweep packed.  On this one-slab ring that is the
 * exchange itself (recv_hi <- send_lo, recv_lo <- send_hi is the identity);
 * a ring of several slabs would move send_hi -> neighbour's recv_lo and
 * send_lo -> the other neighbour's recv_hi between single steps. */
int lb_halo_buffers(lb_ctx *ctx, void **send_lo, void **send_hi, void **recv_lo,
                    void **recv_hi);

/* ---------------------------------------------------------------------- */
/* 3. Batched ragged D2Q9 channels                                          */
/* ---------------------------------------------------------------------- */
typedef struct {
    int nx, ny;
    float tau_inv;
    float u_in;             /* Zou-He inlet velocity at x = 0 */
    int64_t solid_offset;   /* offset of this lattice's mask in the packed mask array */
} lb_batch_lattice;
typedef struct lb_batch lb_batch;
/* Walls (half-way bounce-back) at rows 0 and ny-1, Zou-He velocity inlet
 * u = (u_in, 0) at x = 0 and Zou-He pressure outlet rho = 1 at x = nx-1 (the
 * reference's inlet/outlet, collision.cu:211-262), rest closure on, FP32
 * deviation storage (LB_STORE_F32S); solid masks packed back to back.
 * lb_batch_get_populations returns the stored deviations f_i - w_i. */
int lb_batch_create(int n, const lb_batch_lattice *lat, const uint8_t *solids, lb_batch **out);
int lb_batch_destroy(lb_batch *b);
/* Start from rest (rho 1, u 0) or from device fields packed like the masks
 * (rho, ux, uy planes per lattice: 3*N_b floats per lattice, back to back). */
int lb_batch_init_rest(lb_batch *b);
int lb_batch_init_fields_device(lb_batch *b, const float *d_fields);
/* Same with a choice of initial populations: mode 0 equilibrium (as above),
 * mode 1 equilibrium plus the first-order Chapman-Enskog non-equilibrium
 * part from the fields' velocity gradients, in its post-collision form
 * -(w_i rho (tau - 1) / cs^2) Q_i:S. */
int lb_batch_init_fields_device_ex(lb_batch *b, const float *d_fields, int mode);
int lb_batch_step(lb_batch *b, int nsteps);
/* Run to residual < tol (checked every k steps) or max_steps; per lattice
 * steps taken and final residual.  Converged lattices retire; a lattice whose
 * |u| exceeds 0.5 or turns non-finite retires with residual +inf.  The loop
 * runs on the device: a captured graph of 20 chunks (k steps, residual,
 * retirement, compaction of the block table, all kernels) is built once and
 * replayed; the host reads the active count after each replay. */
int lb_batch_run_to_steady(lb_batch *b, double tol, int k, int max_steps, int *steps_out,
                           double *res_out);
/* Macros (rho, ux, uy) of lattice i to host. */
int lb_batch_get_macros(lb_batch *b, int i, float *rho, float *ux, float *uy);
int lb_batch_get_populations(lb_batch *b, int i, float *host);
int64_t lb_batch_total_nodes(lb_batch *b);

/* ---------------------------------------------------------------------- */
/* 4. U-Net surrogate (steady (ux, uy, rho) from mask, u_in, tau)           */
/*    No counterpart in the reference: the learned warm start.              */
/* ---------------------------------------------------------------------- */
typedef struct {
    int levels;             /* resolution levels (4), 2..6 */
    int base;               /* channels at level 0 (32); doubles per level;
                               a multiple of 32, <= 256, base << (levels-1) <= 1024 */
    int in_ch, out_ch;      /* 3 -> 3; in_ch <= 16, out_ch <= 8 */
    int height, width;      /* 128 x 128 */
    int max_batch;
} lb_unet_desc;
typedef struct lb_unet lb_unet;
/* He-uniform weights from SplitMix64(seed).  Parameter vector layout, in
 * forward conv order: W[co][tap][ci] (ci padded to 16 for the first conv),
 * then bias[co]; finally the 1x1 head W[out][base], bias[out]. */
int lb_unet_create(const lb_unet_desc *desc, uint64_t seed, lb_unet **out);
int lb_unet_destroy(lb_unet *u);
int64_t lb_unet_param_count(lb_unet *u);
int lb_unet_get_params(lb_unet *u, float *host_params, float *host_grads);
int lb_unet_set_params(lb_unet *u, const float *host_params);
/* x: device (B, H, W, in_ch) FP32; y: device (B, H, W, out_ch) FP32 (may be
 * NULL: the output then stays inside the handle for lb_unet_backward and the
 * call is asynchronous on the handle's stream).  With y the call returns
 * once y is written. */
int lb_unet_forward(lb_unet *u, const float *d_x, int B, float *d_y);
/* Masked MSE of the last forward vs d_t (B,H,W,out_ch) with weights d_mask
 * (B,H,W); fills every parameter gradient (asynchronous on the handle's
 * stream; an all-zero mask gives zero gradients).  With loss non-NULL the
 * call waits for the stream and sets *loss = mean squared error
 * (LB_ERR_INVALID for an all-zero mask). */
int lb_unet_backward(lb_unet *u, const float *d_t, const float *d_mask, int B, double *loss);
int lb_unet_adam(lb_unet *u, float lr, float beta1, float beta2, float eps);
int lb_unet_sync(lb_unet *u);
/* Warm-start fields from a standardised prediction d_y (B,H,W,3): per sample
 * b the planes [rho | ux | uy] at d_fields + 3*H*W*b, with
 * ux = u_in (y0 s0 + m0), uy = u_in (y1 s1 + m1), rho = 1 + u_in (y2 s2 + m2),
 * stats6 = {m0, s0, m1, s1, m2, s2} (host), d_uin = u_in per sample (device). */
int lb_unet_to_fields(lb_unet *u, const float *d_y, int B, const float *d_uin, const float *stats6,
                      float *d_fields);
/* One 3x3 convolution (pad 1) on device NHWC tensors, BF16 matrix cores:
 * mode 0: BF16 in/out, bias + ReLU; mode 1: FP32 in/out, no bias.
 * Modes 2 / 3: the same on the per-wavefront gathering kernel only (0 / 1
 * take the LDS-tile kernel when Ci % 32 == 0 and W % 32 == 0, H % 4 == 0 or
 * W % 16 == 0, H % 8 == 0; a K-split tile launch allocates its scratch and
 * the call then waits for the stream).
 * W is BF16 [Co][9][Ci]; Ci % 16 == 0, Co % 32 == 0. */
int lb_conv3x3_bf16(int mode, const void *d_in, const void *d_w, const float *d_bias, void *d_out, int B, int H,
                    int W, int Ci, int Co, lb_stream_t stream);
/* lb_conv3x3_bf16 without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under hipGraph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels
 * and results.  A K-split launch keeps its shares in a fixed per-GPU device
 * array of 16 MiB (4 Mi floats: e.g. batch 16 of 16x16 images with 256
 * output channels split 4 ways); a call needing more returns LB_ERR_SCRATCH
 * and launches nothing.  Asynchronous on `stream` (it never waits).  Every
 * call on one GPU uses that GPU's array: calls that overlap on one GPU (other
 * streams, other threads) race; calls ordered on one stream do not. */
int lb_conv3x3_bf16_noalloc(int mode, const void *d_in, const void *d_w, const float *d_bias, void *d_out, int B,
                            int H, int W, int Ci, int Co, lb_stream_t stream);
/* Weight gradient of one 3x3 convolution (pad 1): d_dw[Co][9][Ci] (FP32) =
 * sum over pixels of d_dout (FP32 NHWC, rounded to BF16 as an MFMA operand)
 * times the shifted input d_in (BF16 NHWC).  method 0: per-element gathers;
 * 1: channel-major transposed operands, 16-byte loads (W % 16 == 0, else
 * LB_ERR_UNSUPPORTED).  Both sum in the same order (identical results).
 * 2: LDS tiles of 8 rows x 32 pixels (W % 32 == 0 and H % 8 == 0) or 16 x
 * 16 (W % 16 == 0 and H % 16 == 0), else LB_ERR_UNSUPPORTED; the kernel the
 * U-Net takes for such shapes; its FP32 summation order differs from 0 / 1.
 * Allocates its scratch and synchronises: a test and diagnostic entry point;
 * lb_unet_backward uses the same kernels on the handle's buffers. */
int lb_conv3x3_wgrad_bf16(int method, const float *d_dout, const void *d_in, float *d_dw, int B, int H, int W,
                          int Ci, int Co, lb_stream_t stream);
/* Diagnostic: rounds every 32-bit pattern to BF16 on the device with the
 * conversion instruction the BF16 stores use and with the bit-pattern
 * round-to-nearest-even of the oracle; *mismatches = non-NaN inputs where
 * the two differ (0 expected).  Synchronises; allocates 8 bytes. */
int lb_check_bf16_rounding(uint64_t *mismatches);
/* Diagnostic: for every 32-bit pattern and two random words derived from
 * (seed, pattern), the stochastic-rounding conversions of the 16-bit stores
 * (round_mode 1: the gfx950 conversion instructions) against the oracle's
 * bit-pattern formulas; mismatches[0] = BF16, mismatches[1] = FP16 pairs
 * that differ (0 expected).  Synchronises; allocates 16 bytes. */
int lb_check_sr_rounding(uint32_t seed, uint64_t *mismatches);

/* ---------------------------------------------------------------------- */
/* 5. Input decks and VTK output (host only)                                */
/* ---------------------------------------------------------------------- */
/* A parameter file (.dat) as Input::loadParameters reads it (input.cpp:51-97):
 * device count, maxDim, ux, uy, uz, rho, viscosity, updateSteps, Fx, Fy, Fz,
 * G, Gads, one per line; lines past the end read as 0.  max_dim <= 1 means a
 * 2D deck whose solids come from the BMP of the same name, > 1 a 3D deck
 * voxelised from the STL of the same name.  tau_inv = 1/(3 nu + 1/2)
 * (input.cpp:168-171); the colour-map bounds follow input.cpp:90-93. */
typedef struct {
    int device_count, max_dim, update_steps;
    float ux, uy, uz, rho, viscosity, tau_inv;
    float fx, fy, fz, G, Gads;
    float min_u, max_u, min_rho, max_rho;
} lb_deck_params;
int lb_deck_read_params(const char *dat_path, lb_deck_params *out);
/* Solids of a 2D deck (replaces Input::loadSolids, input.cpp:99-116): black
 * pixels are solid, solids[x + width*y] with y = 0 the top row.  With
 * solids == NULL only the size is returned.  Uncompressed BMP, 1/4/8/24/32
 * bits per pixel. */
int lb_deck_read_bmp(const char *bmp_path, int *width, int *height, uint8_t *solids);
/* Solids of a 3D deck (replaces Input::loadSolidsSTL, input.cpp:118-142): the
 * mesh (binary or ASCII STL) is scaled so its longest extent spans max_dim
 * voxels; a voxel is solid when its centre is inside the closed surface.
 * solids[x + width*(y + height*z)]; NULL returns the size only. */
int lb_deck_voxelize_stl(const char *stl_path, int max_dim, int *width, int *height, int *depth,
                         uint8_t *solids);
/* VTK image data in the reference's layout (Output::saveParVTK / saveVTK,
 * output.cpp:52-155): dir/r<id>.pvti indexing device_count pieces
 * dir/r<id>p<dev>.vti, each with rho ("Pressure") and the velocity. */
int lb_write_pvti(const char *dir, int vtk_id, int device_count, int width, int height, int length);
int lb_write_vti(const char *dir, int vtk_id, int device_id, int width, int height, int length,
                 const float *ux, const float *uy, const float *uz, const float *rho);

#ifdef __cplusplus
}

/* C++ callers: the reference's call sites (simulation_d2q9.cu:462-554,
 * simulation_d3q19.cu:352-430) compile unchanged.  The launchers also accept
 * any dim3-like grid / block type (HIP's dim3: unsigned x, y, z), converted
 * field by field, and copy_border_devices_* take the stream by reference as
 * copy_border.h:30-47 declares it.  These are inline overloads of the
 * extern "C" entry points above; the library exports only those. */
template <class D3> inline lb_dim3 lb_dim3_of(const D3 &d) {
    lb_dim3 r = {(unsigned int)d.x, (unsigned int)d.y, (unsigned int)d.z};
    return r;
}
template <class D3> inline void streaming_d2q9(DataD2Q9 *data, D3 grid, D3 block) {
    streaming_d2q9(data, lb_dim3_of(grid), lb_dim3_of(block));
}
template <class D3> inline void streaming_d3q19(DataD3Q19 *data, D3 grid, D3 block) {
    streaming_d3q19(data, lb_dim3_of(grid), lb_dim3_of(block));
}
template <class D3>
inline void collision_d2q9(float *plot_data, DataD2Q9 *data, float w0, float w1, float w5, float tau_inv,
                           float ux_in, float uy_in, float rho_out, float u_min, float u_max, float rho_min,
                           float rho_max, D3 grid, D3 block, float Fx, float Fy, float G, float Gads) {
    collision_d2q9(plot_data, data, w0, w1, w5, tau_inv, ux_in, uy_in, rho_out, u_min, u_max, rho_min, rho_max,
                   lb_dim3_of(grid), lb_dim3_of(block), Fx, Fy, G, Gads);
}
template <class D3>
inline void collision_d3q19(float *plot_data, DataD3Q19 *data, float w0, float w1, float w7, float tau_inv,
                            float ux_in, float uy_in, float uz_in, float rho_out, float u_min, float u_max,
                            float rho_min, float rho_max, D3 grid, D3 block) {
    collision_d3q19(plot_data, data, w0, w1, w7, tau_inv, ux_in, uy_in, uz_in, rho_out, u_min, u_max, rho_min,
                    rho_max, lb_dim3_of(grid), lb_dim3_of(block));
}
template <class D3> inline void get_border_d2q9(DataD2Q9 *data, D3 grid, D3 block) {
    get_border_d2q9(data, lb_dim3_of(grid), lb_dim3_of(block));
}
template <class D3> inline void get_border_d3q19(DataD3Q19 *data, D3 grid, D3 block) {
    get_border_d3q19(data, lb_dim3_of(grid), lb_dim3_of(block));
}
template <class D3> inline void apply_border_d2q9(DataD2Q9 *data, D3 grid, D3 block) {
    apply_border_d2q9(data, lb_dim3_of(grid), lb_dim3_of(block));
}
template <class D3> inline void apply_border_d3q19(DataD3Q19 *data, D3 grid, D3 block) {
    apply_border_d3q19(data, lb_dim3_of(grid), lb_dim3_of(block));
}
inline void copy_border_devices_d2q9(float *nb_border_f1, float *nb_border_f5, float *nb_border_f8,
                                     float *nb_border_f3, float *nb_border_f6, float *nb_border_f7,
                                     float *border_f1, float *border_f5, float *border_f8, float *border_f3,
                                     float *border_f6, float *border_f7, int height, lb_stream_t &s) {
    copy_border_devices_d2q9(nb_border_f1, nb_border_f5, nb_border_f8, nb_border_f3, nb_border_f6, nb_border_f7,
                             border_f1, border_f5, border_f8, border_f3, border_f6, border_f7, height, &s);
}
inline void copy_border_devices_d3q19(float *nb_border_f1, float *nb_border_f7, float *nb_border_f8,
                                      float *nb_border_f9, float *nb_border_f10, float *nb_border_f2,
                                      float *nb_border_f11, float *nb_border_f12, float *nb_border_f13,
                                      float *nb_border_f14, float *border_f1, float *border_f7,
                                      float *border_f8, float *border_f9, float *border_f10, float *border_f2,
                                      float *border_f11, float *border_f12, float *border_f13,
                                      float *border_f14, int height, int depth, lb_stream_t &s) {
    copy_border_devices_d3q19(nb_border_f1, nb_border_f7, nb_border_f8, nb_border_f9, nb_border_f10, nb_border_f2,
                              nb_border_f11, nb_border_f12, nb_border_f13, nb_border_f14, border_f1, border_f7,
                              border_f8, border_f9, border_f10, border_f2, border_f11, border_f12, border_f13,
                              border_f14, height, depth, &s);
}
#endif
#endif /* LATTIBOL_AMD_H */
