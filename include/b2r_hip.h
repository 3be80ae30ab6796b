/*
 * b2r_hip.h -- C ABI of the MI355X (gfx950) B2R stencil library.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns 0 on
 * success or a negative B2R_E* code (never exits), and records a message that
 * b2r_last_error() returns.  A context owns its device memory and one HIP stream;
 * the caller owns host memory.  A context is not thread-safe.
 *
 * Geometry (the B2R scheme of the reference, b2r.cpp:47-67, b2r.h:72-87):
 *   the global owned grid env_x*env_y*env_z is split into
 *   block_dim_x*block_dim_y*block_dim_z subdomain blocks of B = env/block_dim
 *   owned cells, each stored with a halo of h = radius*R cells per side, so a
 *   block is S = B + 2h cells per axis.  All blocks of a field live on one GPU in
 *   one allocation, block b = bz*BDx*BDy + by*BDx + bx (b2r_comm.hpp:25-28).
 *   Within a block, cell (x,y,z) with 0 <= x < S_x ... sits at element
 *       b*block_stride + x_offset + z*slice + y*pitch + x
 *   (x fastest, as gg_id in heat-3D.cpp:42-45).  For 4- and 8-byte cells
 *   x_offset = 0 and pitch = S_x, so one block is exactly the reference's dense
 *   S_x*S_y*S_z grid; 1-byte cells get a row offset and pitch that put the first
 *   owned cell on a 16-byte boundary.
 *
 * Boundary kinds at the global faces:
 *   B2R_BOUNDARY_FROZEN   halo cells outside the global grid are never written
 *                         (Dirichlet frame; the edge-aware semantics of
 *                         fdm_heat_diffuse_delta1_verify, heat-3D.cu:57-158);
 *   B2R_BOUNDARY_PERIODIC global faces wrap (the _PERIODIC_ switch of
 *                         b2r_config.h:12, unimplemented in the reference).
 */
#ifndef B2R_HIP_H
#define B2R_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ------------------------------------------------------- */
#define B2R_OK            0
#define B2R_E_ARG        -1   /* bad argument (null pointer, bad value)       */
#define B2R_E_GEOM       -2   /* geometry check failed (divisibility, R*D)   */
#define B2R_E_HIP        -3   /* a HIP runtime call failed                   */
#define B2R_E_NOMEM      -4   /* device allocation failed                    */
#define B2R_E_UNSUPPORTED -5  /* radius / mode / dtype not supported          */

#define B2R_BOUNDARY_FROZEN   0
#define B2R_BOUNDARY_PERIODIC 1

/* Heat arithmetic modes.
 * B2R_HEAT_F32: u' = u + ((kx*Lx + ky*Ly) + kz*Lz) in FP32, L_a the radius-D
 *   second difference evaluated outermost tap first (heat-3D.cu:314-318 order).
 * B2R_HEAT_REF: the reference expression with its FP64 promotion (its double
 *   constants promote the sum):
 *   u' = (float)((double)u + adt*((Lx/hx2 + Ly/hy2) + Lz/hz2)), L_a in FP32
 *   exactly as heat-3D.cpp:57-60 (D=1) and heat-3D.cu:314-318, 409-413, 513-517. */
#define B2R_HEAT_F32 0
#define B2R_HEAT_REF 1

typedef struct b2r_ctx b2r_ctx;

typedef struct b2r_geom {
    int env_x, env_y, env_z;                     /* ENV_DIM_X/Y/Z (b2r_config.h:6-8)  */
    int block_dim_x, block_dim_y, block_dim_z;   /* BLOCK_DIM_X/Y/Z (b2r.cpp:140-143) */
    int R;                                       /* B2R_R (b2r.cpp:139)               */
    int radius;                                  /* per-step radius: B2R_D for heat,
                                                    1 for GoL, 4 for elastic FD4      */
    int boundary;                                /* B2R_BOUNDARY_*                    */
    int elem_size;                               /* bytes per cell: 1, 4 or 8         */
} b2r_geom;

typedef struct b2r_layout {
    int B[3];              /* owned cells per block and axis (B2R_B_X/Y/Z)            */
    int S[3];              /* padded block size (B2R_BLOCK_SIZE_X/Y/Z)               */
    int halo;              /* h = radius*R                                            */
    int nblocks;           /* block_dim_x*block_dim_y*block_dim_z                     */
    int x_offset;          /* elements before x = 0 in each row                       */
    int pitch;             /* elements per row                                        */
    long long slice;       /* elements per z plane (pitch*S_y)                        */
    long long block_stride;/* elements per block                                      */
    long long field_elems; /* elements per field (nblocks*block_stride)               */
    int elem_size;
} b2r_layout;

typedef struct b2r_heat_params {
    int   D;           /* stencil radius 1..4: fdm_heat_diffuse_delta1..4            */
    int   mode;        /* B2R_HEAT_F32 or B2R_HEAT_REF                               */
    float k[3];        /* F32 mode: alpha*dt/h_a^2 per axis                          */
    double adt;        /* REF mode: Alpha*Dt (heat-3D.cu:9-10)                      */
    double h2[3];      /* REF mode: Dx*Dx, Dy*Dy, Dz*Dz (heat-3D.cu:11-13)          */
} b2r_heat_params;

/* Launch shapes of a context; 0 selects the built-in automatic choice.  No
 * counterpart in the reference (CUDA_BLOCK_SIZE is a compile-time 32 or 16,
 * heat-3D.cu:8, eqwp.cu:25); set by hand or by b2r_autotune. */
typedef struct b2r_tuning {
    int heat_ry;      /* rows per thread of the heat z-march: 1, 2, 4 or 8         */
    int heat_zc;      /* z planes marched per heat workgroup                        */
    int gol_zc;       /* z planes marched per fused Game of Life workgroup          */
    int gol_step_zc;  /* z planes marched per single-step Game of Life workgroup    */
    int el_zc;        /* z planes marched per elastic workgroup                     */
    int heat_wy;      /* waves (thread rows) per heat step workgroup: 2, 3, 4, 8,
                         16                                                    */
    int gol_tile;     /* fused Game of Life kernel and tile (cells x rows incl. halo
                         rows): 1-3 byte-lane 128 x 32, 128 x 48, 256 x 56; 4-8
                         bit-sliced 256 x 32, 256 x 64, 256 x 96, 256 x 40,
                         256 x 72; 9, 10 bit-sliced 256 x 72, 256 x 136 with
                         two rows per thread; 11, 12 bit-sliced 256 x 72,
                         256 x 48 with waves of whole rows (x neighbours by
                         lane shifts); 13-15 bit-sliced whole grid rows of
                         512 cells x 72 (two rows per thread), x 40, x 64,
                         where the grid is periodic in x over 512 cells (x
                         neighbours by rotations within 16 lanes; elsewhere
                         they run as 11); 0 = 5                                */
    int heat_wx;      /* waves across x per heat step workgroup; (heat_wy,
                         heat_wx) is (4, 1|2|3|4), (8, 1|2), (16, 1), (2, 6) or
                         (3, 3)
                         (heat_wy 0 counts as 4, heat_wx 0 as 1)               */
    int el_kernel;    /* batched elastic step: 1 one fused stress + velocity launch
                         from a second copy of the state (ping-pong), 2 separate
                         stress and velocity launches in place; 0 = 1 where every
                         global face is frozen or all are periodic, else 2       */
    int el_step_zc;   /* z planes marched per fused elastic step workgroup (el_zc
                         serves the separate kernels)                           */
    int el_grad_mb;   /* device memory the batched gradient may take for its
                         checkpoints and adjoint state, MB; the shots run in
                         groups that fit; 0 = the free memory less 5 %          */
    int gol_bytes;    /* fused bit-sliced Game of Life runs with B_x % 32 == 0:
                         0 keep the fields bit-packed (1 bit per cell) between
                         launches, packed at the start of b2r_gol_run and
                         unpacked at its end; 1 byte fields throughout        */
    int gol_exchange; /* 1: fused Game of Life runs exchange the halos with
                         b2r_boundary_updating (pack/unpack) before every
                         R-generation launch, on byte fields, and the kernel
                         reads its own block's halo (the reference's scheme,
                         b2r.cpp:179-221); 0: the bit-sliced kernel reads its
                         halo from the neighbouring blocks, no exchange       */
} b2r_tuning;

/* ---- library / context ---------------------------------------------------- */
const char* b2r_last_error(void);
int b2r_version(void);

/* Validates and creates: the checks of b2r.cpp:154-167 and simu_env_init
 * (b2r.cpp:54-59: radius*R <= min(B)/2), then the per-field layout.  Replaces the
 * geometry globals of b2r.h:72-87 and b2r_sync_init (b2r_halo_sync.cu:74-111). */
int b2r_create(const b2r_geom* geom, int device, b2r_ctx** out);
int b2r_destroy(b2r_ctx* ctx);  /* b2r_sync_finalize (b2r_halo_sync.cu:121-128) */
int b2r_get_layout(const b2r_ctx* ctx, b2r_layout* out);
int b2r_synchronize(b2r_ctx* ctx);
int b2r_set_tuning(b2r_ctx* ctx, const b2r_tuning* t);
int b2r_get_tuning(const b2r_ctx* ctx, b2r_tuning* t);

/* Which global faces of a FROZEN grid are frozen: bit 2*axis+side (axis 0..2 =
 * x,y,z; side 0 = minus, 1 = plus).  Default 0x3F (all).  A cleared bit makes
 * that face open: the blocks on it update their halo there like an interior
 * one, as the reference's general kernels do (heat-3D.cu:173-235, gol-3D.cu:50),
 * and nothing refreshes it; the edge flags ef of heat_3D_gpu_main /
 * fdm_heat_diffuse_delta1_verify (heat-3D.cu:57-114, 583) map to it (ef=-1:
 * minus face frozen, ef=1: plus face frozen, ef=0: both open). */
int b2r_set_frozen_faces(b2r_ctx* ctx, int mask);

/* Neighbour-exists flags of one block, order left,right,up,down,infront,behind:
 * 0 = none (global face), 1 = neighbour, -1 = periodic wrap (b2r.h:37).
 * Replaces b2r_sync_flag(int rank) (b2r_halo_sync.cu:140-183). */
int b2r_sync_flag(const b2r_ctx* ctx, int rank, signed char flags[6]);

/* ---- fields (device memory owned by the context) -------------------------- */
int b2r_field_alloc(b2r_ctx* ctx, void** d_field);
int b2r_field_free(b2r_ctx* ctx, void* d_field);
/* Sets every cell of every block (halos included) to *value (elem_size bytes). */
int b2r_field_fill(b2r_ctx* ctx, void* d_field, const void* value);
/* Copies the global owned grid env_x*env_y*env_z (x fastest) in or out of the
 * owned regions of all blocks; halos are untouched. */
int b2r_field_upload(b2r_ctx* ctx, void* d_field, const void* h_global);
int b2r_field_download(b2r_ctx* ctx, const void* d_field, void* h_global);
/* Raw copy of one padded block as a dense S_x*S_y*S_z host grid, the h_in of
 * heat_3D_gpu_main / gol_3D_gpu_main (heat-3D.cu:583, gol-3D.cu:119). */
int b2r_block_upload(b2r_ctx* ctx, void* d_field, int block, const void* h_block);
int b2r_block_download(b2r_ctx* ctx, const void* d_field, int block, void* h_block);
int b2r_field_copy(b2r_ctx* ctx, void* d_dst, const void* d_src);

/* Seeded synthetic inputs, keyed by the global cell index gid = x+ENVx*(y+ENVy*z)
 * through SplitMix64 (identical on CPU and GPU):
 *   uniform:   f32 cells u = (splitmix64(splitmix64(seed)+gid) >> 40) * 2^-24
 *   bernoulli: u8 cells  (splitmix64(splitmix64(seed)+gid) >> 40) < p_q24
 * Owned cells only; halos untouched. */
int b2r_field_init_uniform(b2r_ctx* ctx, void* d_field, uint64_t seed);
int b2r_field_init_bernoulli(b2r_ctx* ctx, void* d_field, uint64_t seed, uint32_t p_q24);
/* Separable sum of nb Gaussian blobs, evaluated in FP64 and rounded to f32:
 * u(x,y,z) = sum_k amp[k]*ex[k*env_x+x]*ey[k*env_y+y]*ez[k*env_z+z], the sum taken
 * k = 0..nb-1 left to right.  ex/ey/ez are host tables of nb*env_* doubles. */
int b2r_field_init_blobs(b2r_ctx* ctx, void* d_field, int nb, const double* amp,
                         const double* ex, const double* ey, const double* ez);

/* ---- halo exchange (boundary_updating) ------------------------------------ */
/* Ordered X -> Y -> Z exchange of the h-deep halos of every block of one field,
 * through device pack/unpack kernels and device message buffers in the message
 * layout of b2r_comm.hpp (X: [x][y][z] over B_Y*h*B_Z, Y: [y][z][x] over
 * S_X*h*B_Z, Z: [z][y][x] over S_X*S_Y*h).  Replaces boundary_updating
 * (b2r_halo_sync.cu:442-508); time_step only labels the exchange (MPI tag). */
int b2r_boundary_updating(b2r_ctx* ctx, int time_step, void* d_field);
/* b2r_boundary_updating without any allocation: the message buffers live in a
 * fixed 64 MiB device array of the library (one per GPU, its address taken
 * once per device by b2r_create).  No counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Fails with
 * B2R_E_ARG when one direction's messages (2 * nblocks * the largest
 * b2r_msg_elems * elem_size bytes) exceed 64 MiB.  Every call on one GPU
 * shares the array: calls that overlap on one GPU (other streams, other
 * contexts, other threads) race.  Result equal to b2r_boundary_updating. */
int b2r_boundary_updating_noalloc(b2r_ctx* ctx, int time_step, void* d_field);
/* Same result with each halo cell copied straight from its source block (no
 * message buffers): the device-direct path of boundary_updating_direct
 * (b2r_halo_sync.cu:389-431).  Several fields in one launch per direction;
 * no launch along an axis with one block and frozen faces (no neighbours). */
int b2r_boundary_updating_direct(b2r_ctx* ctx, int time_step, void* const* d_fields, int nfields);
/* Message-level building blocks: pack (sub_matrix_pack_{x,y,z},
 * b2r_halo_sync.cu:195-282) every block's two faces along dim (0,1,2) into
 * d_msgs, and unpack (sub_matrix_unpack_{x,y,z}, :292-379) the messages
 * addressed to each block.  d_msgs holds nblocks*2 messages of
 * b2r_msg_elems(dim) elements: message (b, side) is what block b sends to its
 * side-0 (minus) or side-1 (plus) neighbour. */
long long b2r_msg_elems(const b2r_ctx* ctx, int dim);
int b2r_pack(b2r_ctx* ctx, int dim, const void* d_field, void* d_msgs);
int b2r_unpack(b2r_ctx* ctx, int dim, void* d_field, const void* d_msgs);

/* ---- heat diffusion (heat-3D.cu) ------------------------------------------ */
/* One inner B2R step b2r_i of fdm_heat_diffuse_delta{D} over every block:
 * out = step(in) on the region [D*(b2r_i+1), S-1-D*(b2r_i+1)] per axis, clipped
 * at frozen global faces to the owned range (heat-3D.cu:173-235, 250-521,
 * 57-158).  Cells outside the region are not written.  geom.radius must be D,
 * and 2D+2 x-y planes of a block must stay below 4 GiB (32-bit offsets). */
int b2r_heat_step(b2r_ctx* ctx, const void* d_in, void* d_out,
                  const b2r_heat_params* p, int b2r_i);
/* nsteps steps of the B2R loop of b2r.cpp:179-221 with heat_3D_gpu_main's R-step
 * ping-pong (heat-3D.cu:602-628): every R steps boundary_updating_direct, then
 * R inner steps.  The result is in d_a if *result_in_b == 0, else in d_b.
 * d_a and d_b must hold identical halos (the frozen frame) on entry.
 * fused != 0 runs the R inner steps of each exchange period in one launch
 * (temporal blocking in LDS) where R*D <= 8 and every face is frozen or
 * periodic; elsewhere it runs them step by step.  *ms receives the device time. */
int b2r_heat_run(b2r_ctx* ctx, void* d_a, void* d_b, const b2r_heat_params* p,
                 int nsteps, int fused, int* result_in_b, float* ms);
/* Hand-written discrete adjoint of the heat time loop; no reference
 * counterpart (the reference runs heat forward only, heat-3D.cu:583-642).
 * J = 1/2 sum over owned cells of (u_N - target)^2 with u_N = nsteps steps of
 * delta-D from d_u0.  Writes dJ/du_0 (owned cells) into d_grad_u0 and
 * dJ/dk_a, a = x, y, z, into grad_k[3] (REF mode: k_a = adt / h2_a).  Needs
 * periodic or frozen global faces.  Forward states are checkpointed every
 * ckpt steps and recomputed per segment: ceil(nsteps/ckpt) + max(ckpt, 2) + 2
 * extra fields of device memory.  FP64 reductions; d_u0 and d_target are
 * read only (they may alias d_grad_u0). */
int b2r_heat_gradient(b2r_ctx* ctx, const void* d_u0, const void* d_target,
                      const b2r_heat_params* p, int nsteps, int ckpt, double* misfit,
                      void* d_grad_u0, double* grad_k);

/* One inner step b2r_i of delta-D on the host cores over one dense padded
 * Ngx*Ngy*Ngz host grid (x fastest): out = step(in) on [D(b2r_i+1),
 * N-1-D(b2r_i+1)] per axis, the arithmetic of the device kernels term for term
 * (either mode), so a host step equals a device step bit for bit.  Replaces
 * cpu_fdm_heat_diffuse1..4 (heat-3D.cpp:48-134; their delta2-4 use the bounds
 * and taps of delta1, which is not reproduced); nthreads <= 0 takes every hardware
 * thread.  No device, no context. */
int b2r_heat_cpu_step(const b2r_heat_params* p, const float* h_in, float* h_out, int Ngx,
                      int Ngy, int Ngz, int b2r_i, int nthreads);
/* b2r_heat_run with the host cores taking part: they own the top cpu_planes
 * owned z planes of the global grid (R*D <= cpu_planes <= env_z - R*D), the
 * GPU the rest; every R steps the two exchange joint layers of R*D planes over
 * PCIe and then take R inner steps at once (the CPU+GPU scheme of the B2R paper:
 * the "joint layer" of heat-3D.cpp:47, 55 and heat_3D_gpu_diffsync_main,
 * declared at heat-3D.hpp:15 and never defined there).  Equal bit for bit to
 * b2r_heat_run without fusion; the slab's planes are written back into the
 * result field.  3D grids with every global face frozen; *ms is wall time. */
int b2r_heat_run_hybrid(b2r_ctx* ctx, void* d_a, void* d_b, const b2r_heat_params* p,
                        int nsteps, int cpu_planes, int nthreads, int* result_in_b, float* ms);

/* ---- 3D Game of Life (gol-3D.cu) ------------------------------------------ */
/* One generation (26-neighbour count n; alive & n<=1 -> 0; dead & n==5 -> 1;
 * alive & n>=8 -> 0; otherwise unchanged; gol-3D.cu:83-99) over the inner-step
 * region of b2r_i (gol-3D.cu:50, 58).  u8 cells in {0,1}; geom.radius must be 1. */
int b2r_gol_step(b2r_ctx* ctx, const void* d_in, void* d_out, int b2r_i);
/* nsteps generations of every block: d_b := d_a first (gol-3D.cu:128-129), then
 * per window of R generations the exchange and R b2r_gol_step launches
 * alternating the two fields (the loop of gol-3D.cu:138-145 plus the driver's
 * exchange, b2r.cpp:179-221), or with fused = 1 (periodic, or every face
 * frozen; R <= 4) one R-generation launch per window.  Where the bit-sliced
 * kernel runs (B_x % 32 == 0, gol_tile 0 or 4-15) no exchange runs: it reads
 * its halo from the neighbouring blocks, and unless gol_bytes is set the
 * windows run on the context's two bit-packed fields (allocated on the first
 * such call), nsteps mod R generations at the end as one more, shorter
 * launch, the result's owned cells then written back to the byte field.
 * *result_in_b says which field holds the result; *ms, when not NULL, the
 * device time between events around the loop. */
int b2r_gol_run(b2r_ctx* ctx, void* d_a, void* d_b, int nsteps, int fused,
                int* result_in_b, float* ms);

/* ---- elastic velocity-stress FD4 (eqwp.cu) ------------------------------------ */
/* One shot: 9 state fields (vx vy vz sxx syy szz sxy sxz syz, index 0..8) and
 * the medium lambda, mu, b = 1/rho (9, 10, 11) on a context of radius 4 (one
 * step = stress stage + velocity stage, radius 2 each, so an R-step window
 * needs a 4R halo).  Replaces eqwp_cuda_init (eqwp.cu:515-582); per step:
 *   sigma += dt*(lambda tr(E) I + 2 mu E),  E = sym grad v  (eqwp_fd4_stress,
 *            eqwp.cu:341-477, lambda+2mu / lambda / mu of :460-466)
 *   v     += dt*b*div(sigma)                (eqwp_fd4_vx/vy/vz, :60-331)
 * first derivatives (-1/12 (f[+2]-f[-2]) + 2/3 (f[+1]-f[-1])) * inv_h (:22-23),
 * then point forces v_c += dt*b*f(t) (also at a source's periodic images in
 * halos) and receiver samples after the step.  A block's field must stay below
 * 4 GiB (B2R_E_UNSUPPORTED otherwise): the stencil kernels address it by
 * 32-bit byte offsets. */
typedef struct b2r_el_shot b2r_el_shot;
int b2r_el_create(b2r_ctx* ctx, float dt, const float inv_h[3], b2r_el_shot** out);
int b2r_el_destroy(b2r_el_shot* s);
int b2r_el_field(const b2r_el_shot* s, int which, void** d_field);
/* medium on the global owned grid (x fastest); halos filled by exchange */
int b2r_el_set_medium(b2r_el_shot* s, const float* lam, const float* mu, const float* b);
int b2r_el_reset(b2r_el_shot* s);  /* zero state, t = 0 */
/* the displacement accumulator of eqwp_fd4_vy (eqwp.cu:229-235):
 * d += dt*dt*b*((div sigma)_y + f_y), f_y the point forces in component 1 */
int b2r_el_enable_displacement(b2r_el_shot* s);
/* nsrc point forces at global cells xyz[3i..], component comp[i] (0..2), values
 * wav[t*nsrc + i] for steps t < nt (the force patch of eqwp.cu:561-566 is one
 * source per cell with a constant value for t < FRC_TIME). */
int b2r_el_set_sources(b2r_el_shot* s, int nsrc, const int* xyz, const int* comp, int nt,
                       const float* wav);
/* nrec receivers sampling field comp (0..8) after each step t < nt */
int b2r_el_set_receivers(b2r_el_shot* s, int nrec, const int* xyz, int comp, int nt);
int b2r_el_get_receivers(b2r_el_shot* s, float* h_rec);           /* nt x nrec */
int b2r_el_get_state(b2r_el_shot* s, int which, float* h_global); /* -1: displacement */
int b2r_el_set_state(b2r_el_shot* s, int which, const float* h_global);
/* nsteps steps of b2r.cpp:179-221 / eqwp_gpu_main (eqwp.cu:661-714): every R
 * steps the 9 fields are exchanged, then R steps.  mode 0 (plain): per shot,
 * stress, vx, vy, vz launches and a device synchronisation each step;
 * mode 1 (batched): one stress and one velocity launch per step for all shots;
 * mode 2: mode 1 with the time loop replayed from HIP graphs.  The shots of a
 * batch may differ in size; they must share the device and R. */
int b2r_el_forward(b2r_el_shot* s, int nsteps, int mode, float* ms);
int b2r_el_forward_batch(b2r_el_shot* const* shots, int nshots, int nsteps, int mode, float* ms);
/* misfit 1/2 sum (rec - obs)^2 over the recorded samples, reduced in FP64 */
int b2r_el_misfit(b2r_el_shot* s, const float* h_obs, double* misfit);
/* Discrete adjoint (no counterpart in the reference): misfit of nsteps steps
 * from a zero state and its gradient w.r.t. lambda, mu and b per owned cell
 * (global grids).  The state is checkpointed every ckpt steps; the reverse
 * sweep undoes each step (the velocity-stress scheme is time-reversible)
 * together with its adjoint and resets the state from the checkpoint at every
 * segment boundary, so nothing is stored between checkpoints:
 * (ceil(nsteps/ckpt) - 1) * 9 + 12 extra fields.  R = 1, every global face
 * frozen or periodic, any block decomposition (the state and the adjoint are
 * exchanged before every reversed step); sources at distinct cells. */
int b2r_el_gradient(b2r_el_shot* s, const float* h_obs, int nsteps, int ckpt, double* misfit,
                    float* h_glam, float* h_gmu, float* h_gb);
/* flags bit 0: start from the current state (not zero); bit 1: the plain
 * algorithm (forward in the reference's per-step structure, each segment
 * recomputed with all its states kept, ckpt + 1 + ceil(nsteps/ckpt) state
 * copies; one block with frozen faces); h_adj0 (optional, 9 global grids):
 * the adjoint state at the start = d misfit / d initial state */
int b2r_el_gradient_ex(b2r_el_shot* s, const float* h_obs, int nsteps, int ckpt, int flags,
                       double* misfit, float* h_glam, float* h_gmu, float* h_gb, float* h_adj0);
/* Gradients of nshots shots (sizes may differ), shot k with observations
 * h_obs[k] and outputs misfit[k], h_glam[k], h_gmu[k], h_gb[k].  mode 0: the
 * plain per-shot loop (flags bit 1 above, one shot after the other); mode 1:
 * all shots per launch, reverse sweep as b2r_el_gradient; mode 2: mode 1 with
 * every checkpoint segment replayed from a HIP graph.  Modes 1 and 2 run the
 * shots in consecutive groups whose checkpoints and adjoint state fit the
 * memory b2r_tuning.el_grad_mb allows (16 shots of 256^3 over 1500 steps with
 * checkpoints every 50 need about 320 GB); a single shot that does not fit is
 * an error.  *ms: device time, summed over the groups. */
int b2r_el_gradient_batch(b2r_el_shot* const* shots, int nshots, const float* const* h_obs,
                          int nsteps, int ckpt, int mode, double* misfit, float* const* h_glam,
                          float* const* h_gmu, float* const* h_gb, float* ms);

/* How b2r_el_gradient_batch (modes 1, 2) would split these shots: the number
 * of consecutive groups whose checkpoints and adjoint state fit the memory
 * b2r_tuning.el_grad_mb allows, and the device bytes of the largest group.
 * No counterpart in the reference. */
int b2r_el_gradient_plan(b2r_el_shot* const* shots, int nshots, int nsteps, int ckpt,
                         int* ngroups, size_t* bytes_max_group);

/* ---- measurement helpers ---------------------------------------------------- */
/* iters device-to-device copies of bytes (a STREAM-copy kernel): *ms total. */
int b2r_stream_copy(b2r_ctx* ctx, void* d_dst, const void* d_src, size_t bytes,
                    int iters, float* ms);
/* iters back-to-back launches of the inner-step kernel (b2r_i = it % R) alternating
 * d_a -> d_b -> d_a ..., bracketed by HIP events on the context stream; *ms is the
 * total device time (kind 0: heat with *p, kind 1: Game of Life).  No counterpart
 * in the reference: it times the kernel alone for the roofline figure. */
int b2r_time_steps(b2r_ctx* ctx, int kind, void* d_a, void* d_b, const b2r_heat_params* p,
                   int iters, float* ms);
/* Autotuner (no counterpart in the reference): times every candidate launch
 * shape of one kernel kind (as b2r_time_steps: 0 heat step, 1 GoL step, 2 fused
 * GoL, 3 fused heat) on this context's geometry, keeps the fastest in the
 * context (b2r_get_tuning) and, if record_path is given, appends one text line
 * with the geometry, the choice and its time per launch. */
int b2r_autotune(b2r_ctx* ctx, int kind, void* d_a, void* d_b, const b2r_heat_params* p,
                 const char* record_path, b2r_tuning* out);
/* The elastic kernels' autotuner (no counterpart in the reference): times the
 * batched step of one shot with the fused step kernel (where the faces allow
 * it) and with the separate stress and velocity kernels, for each z-chunk
 * candidate, keeps each kernel's fastest z chunk (b2r_tuning.el_step_zc,
 * el_zc) and the faster kernel (el_kernel) in the shot's context, appends a
 * record line as above, and leaves the shot reset (zero state, t = 0). */
int b2r_el_autotune(b2r_el_shot* s, const char* record_path, b2r_tuning* out);
/* The same for a batch as b2r_el_forward_batch runs it (no counterpart in the
 * reference): each candidate (kernel, z chunk) is set in every shot's context
 * and timed on the whole batch, whose many workgroups load the GPU
 * differently from any one shot; the fastest is kept in every context, one
 * record line is appended, and the shots are left reset. */
int b2r_el_autotune_batch(b2r_el_shot* const* shots, int nshots, const char* record_path);
/* Scratch device buffer owned by the context (freed by b2r_destroy). */
int b2r_dev_alloc(b2r_ctx* ctx, size_t bytes, void** d_ptr);
int b2r_dev_free(b2r_ctx* ctx, void* d_ptr);
int b2r_memcpy_h2d(b2r_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int b2r_memcpy_d2h(b2r_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* B2R_HIP_H */
