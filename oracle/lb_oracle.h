/*
 * lb_oracle.h -- CPU restatement of the lattibol lattice-Boltzmann hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product library never links it.
 *
 * Two families of functions:
 *
 *  orc_ref_*   A transcription of the reference CUDA kernels and host drivers
 *              with the reference's arithmetic order and float/double
 *              promotions (in-place "swap" streaming, full-way bounce-back,
 *              Zou-He inlet/outlet, D2Q9 Shan-Chen with adhesion, halo
 *              pack/unpack, point probe).  FP32 only, like the reference.
 *
 *  orc_pull_*  The specification of the fused pull-scheme sweep the GPU path
 *              implements (A/B buffers, stream+BC+force+BGK per node),
 *              templated over the storage type (FP32, FP64, FP16/BF16 shifted
 *              deviations f_i - w_i).  Its BGK, moments and Zou-He keep the
 *              reference arithmetic order, so with full-way bounce-back and
 *              Zou-He switched on it is bit-identical to orc_ref_* on fluid
 *              nodes (that equivalence is itself a test).
 *
 * Population layout everywhere: structure of arrays, population i of node
 * n = x + nx*(y + ny*z) at index i*N + n, N = nx*ny*nz.
 *
 * Pinning (which independent check covers which function) is listed in the
 * header comment of lb_oracle.cpp.
 */
#ifndef LB_ORACLE_H
#define LB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary kinds per face; faces are 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z in storage
 * index space (D2Q9 rows grow downward, physical y = -row). */
enum { ORC_BC_PERIODIC = 0, ORC_BC_WALL = 1, ORC_BC_MOVING_WALL = 2,
       ORC_BC_INLET_EQ = 3, ORC_BC_OUTFLOW = 4 };
/* F32/F64 store f_i (reference expressions); F16/BF16/F32S store
 * s_i = f_i - w_i and collide in deviation form. */
enum { ORC_STORE_F32 = 0, ORC_STORE_F16 = 1, ORC_STORE_BF16 = 2, ORC_STORE_F64 = 3, ORC_STORE_F32S = 4 };

typedef struct {
    int lattice;            /* 9 (D2Q9) or 19 (D3Q19) */
    int nx, ny, nz;         /* nz = 1 for D2Q9 */
    int storage;            /* ORC_STORE_* */
    float tau_inv;          /* 1/tau */
    int bc[6];              /* ORC_BC_* per face */
    float bc_u[6][3];       /* physical wall / inlet velocity per face */
    float bc_rho[6];        /* inlet density per face */
    float force[3];         /* body force (velocity-shift forcing) */
    float sc_G;             /* Shan-Chen coupling; 0 disables */
    float sc_rho0;          /* psi = rho0 * (1 - exp(-rho/rho0)) */
    float sc_psi_solid;     /* psi assigned to solid nodes */
    int bb_full;            /* 1: full-way bounce-back at solid nodes */
    int zou_he;             /* 1: reference Zou-He inlet (x=0) / outlet (x=nx-1) */
    float zh_ux_in, zh_rho_out;
    int rest_closure;       /* 1: feq_0 = rho - sum_{i>0} feq_i */
    int round_mode;         /* 16-bit stores: 0 nearest-even, 1 stochastic (orc_sr_bits) */
    uint32_t step;          /* steps since initialisation (the stochastic-rounding stream) */
} orc_cfg;

/* ---- shared numerics -------------------------------------------------- */
double   orc_exp(double x);
float    orc_expf(float x);
float    orc_psi_f32(float rho, float rho0);
double   orc_psi_f64(double rho, float rho0);
uint64_t orc_splitmix64_next(uint64_t *state);
void     orc_uniform_noise(uint64_t seed, int64_t n, double base, double amp, float *out);
uint16_t orc_f32_to_f16(float f);
float    orc_f16_to_f32(uint16_t h);
uint16_t orc_f32_to_bf16(float f);
/* Stochastic rounding: 32 random bits of population slot i of node id at
 * time step `step`, and the rounding of v to FP16 / BF16 with them. */
uint32_t orc_sr_bits(uint32_t step, uint32_t id, int slot);
uint16_t orc_f32_to_f16_sr(float f, uint32_t r);
uint16_t orc_f32_to_bf16_sr(float f, uint32_t r);
float    orc_bf16_to_f32(uint16_t h);
void     orc_f32_to_f16_array(const float *in, uint16_t *out, int64_t n);
void     orc_f32_to_bf16_array(const float *in, uint16_t *out, int64_t n);
void     orc_set_threads(int n);
int      orc_get_threads(void);

/* ---- reference transcription (FP32) ------------------------------------ */
void orc_ref_streaming_d2q9(float *f, const uint8_t *solid_bound, int W, int H);
void orc_ref_streaming_d3q19(float *f, const uint8_t *solid_bound, int W, int H, int D);
void orc_ref_collision_d2q9(float *plot, float *f, const uint8_t *solid, float *ux, float *uy,
                            float *rho, float *psi, float *ueq_x, float *ueq_y, int W, int H,
                            int dev, int ndev, float w0, float w1, float w5, float tau_inv,
                            float ux_in, float uy_in, float rho_out, float Fx, float Fy,
                            float G, float Gads, int zou_he);
void orc_ref_collision_d3q19(float *plot, float *f, const uint8_t *solid, float *ux, float *uy,
                             float *uz, float *rho, int W, int H, int D, int dev, int ndev,
                             float w0, float w1, float w7, float tau_inv, float ux_in,
                             float uy_in, float uz_in, float rho_out, int zou_he);
void orc_ref_get_border_d2q9(const float *f, float *border6, int W, int H);
void orc_ref_apply_border_d2q9(float *f, const float *nb6, int W, int H);
void orc_ref_get_border_d3q19(const float *f, float *border10, int W, int H, int D);
void orc_ref_apply_border_d3q19(float *f, const float *nb10, int W, int H, int D);
void orc_ref_read_value(const float *plot, int W, int H, int x, int y, int z, float out[4]);
void orc_ref_solid_bound_d2q9(const uint8_t *solid, uint8_t *sb, int W, int H);
void orc_ref_solid_bound_d3q19(const uint8_t *solid, uint8_t *sb, int W, int H, int D);
/* Host initialisation of every subdomain (arrays hold ndev consecutive
 * subdomains of W_sub*H(*D) nodes).  Returns 0, or -1 for Shan-Chen with
 * ndev > 1 (the reference initialiser indexes past its buffers there). */
int orc_ref_init_d2q9(const uint8_t *all_solids, int Wfull, int H, int ndev, float ux0,
                      float uy0, float rho0, float tau_inv, float Fx, float Fy, float G,
                      float Gads, float *f, float *ux, float *uy, float *rho, float *psi,
                      float *ueq_x, float *ueq_y, uint8_t *solid, uint8_t *solid_bound);
int orc_ref_init_d3q19(const uint8_t *all_solids, int Wfull, int H, int D, int ndev, float ux0,
                       float uy0, float uz0, float rho0, float *f, float *ux, float *uy,
                       float *uz, float *rho, uint8_t *solid, uint8_t *solid_bound);
/* The reference step loop (runLBM) over ndev subdomains on a periodic ring. */
void orc_ref_run_d2q9(int ndev, int W, int H, float *f, const uint8_t *solid,
                      const uint8_t *solid_bound, float *ux, float *uy, float *rho, float *psi,
                      float *ueq_x, float *ueq_y, float *plot, float ux_in, float uy_in,
                      float rho_out, float tau_inv, float Fx, float Fy, float G, float Gads,
                      int steps, int cx, int cy, float value_cur[4], int zou_he);
void orc_ref_run_d3q19(int ndev, int W, int H, int D, float *f, const uint8_t *solid,
                       const uint8_t *solid_bound, float *ux, float *uy, float *uz, float *rho,
                       float *plot, float ux_in, float uy_in, float uz_in, float rho_out,
                       float tau_inv, int steps, int cx, int cy, int cz, float value_cur[4],
                       int zou_he);

/* ---- pull-scheme specification ----------------------------------------- */
/* Equilibrium from fields (double inputs are first rounded to the compute
 * type: float for FP32/FP16/BF16 storage, double for FP64). */
void orc_pull_init_eq(const orc_cfg *cfg, const double *rho, const double *ux,
                      const double *uy, const double *uz, void *A);
/* One time step A -> B.  macros (nullable) receives rho,ux,uy,uz planes in
 * the compute type (4*N entries). */
void orc_pull_step(const orc_cfg *cfg, const void *A, void *B, const uint8_t *solid,
                   void *macros);
/* steps time steps; the final state is left in A. */
void orc_pull_run(orc_cfg *cfg, void *A, void *B, const uint8_t *solid, int steps,
                  void *macros);
/* Decode stored populations to float64 (adds w_i back for shifted storage). */
void orc_pull_decode(const orc_cfg *cfg, const void *A, double *out);
/* FP64 totals over fluid nodes: mass, momentum x, y, z (physical axes). */
void orc_pull_reduce(const orc_cfg *cfg, const void *A, const uint8_t *solid, double out[4]);
/* ||u_now - u_prev||_2 / ||u_now||_2 over fluid nodes; u given as 3 planes
 * of N floats (or 2 for D2Q9). */
double orc_residual(int ncomp, int64_t n, const float *u_now, const float *u_prev,
                    const uint8_t *solid);

#ifdef __cplusplus
}
#endif
#endif
