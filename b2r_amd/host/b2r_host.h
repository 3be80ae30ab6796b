// b2r_host.h -- the reference's host-side interface, kept as thin C++ shims over
// the C ABI (include/b2r_hip.h).  Names, arguments and error behaviour follow
// b2r.cpp, b2r.h, b2r_halo_sync.cu, heat-3D.cu, gol-3D.cu and eqwp.cu: the shims
// print the library's message and exit(-1) on failure, as cudaErrchk does
// (b2r_halo_sync.cu:22-28).  One process drives all Bx*By*Bz blocks on one GPU,
// so "rank" means block index and the fields are device-resident multi-block
// fields of the library.
//
// The globals below are the ones b2r.h:51-87 defines.  libb2r_host.so defines
// them too, so that b2r_main links alone; a driver that includes the
// reference's b2r.h (b2r.cpp) defines its own, and the library's references
// bind to the driver's copies (ELF symbol interposition: the executable's
// definition comes first), which is how the shims read the geometry that the
// driver's simu_env_init computed and hand back d_vx .. d_sigma_zz.
#pragma once

#include "../../include/b2r_hip.h"

typedef float CELL_DT;  // b2r_config.h:17

// geometry globals of b2r.h:72-87, set by the driver's simu_env_init
// (b2r.cpp:47-67) before b2r_sync_init reads them
extern int B2R_B_X, B2R_B_Y, B2R_B_Z;
extern int BLOCK_DIM_X, BLOCK_DIM_Y, BLOCK_DIM_Z;
extern int B2R_BLOCK_SIZE_X, B2R_BLOCK_SIZE_Y, B2R_BLOCK_SIZE_Z;
extern int B2R_R;
extern char b2r_halo_update[6];  // b2r.h:37-42 (4 entries in 2D)

// device fields of b2r.h:60-64; eqwp_cuda_init points them at the elastic
// shot's fields (library fields holding every block)
extern CELL_DT *d_vx, *d_vy, *d_vz;
extern CELL_DT *d_sigma_xx, *d_sigma_xy, *d_sigma_xz, *d_sigma_yy, *d_sigma_yz, *d_sigma_zz;

// Compile-time switches of the reference's b2r_config.h:6-13.  Built with
// CONFIG=path/to/b2r_config.h (b2r_amd/host/Makefile) the library takes them
// from that header; otherwise ENV 256^3, B2R_D 1, _PERIODIC_ 0, _ENV_3D_ 1.
// b2r_main sets ENV_DIM_* and B2R_D from its command line.
extern int ENV_DIM_X, ENV_DIM_Y, ENV_DIM_Z;  // ENV_DIM_X/Y/Z
extern int B2R_D;                            // B2R_D (heat/gol model entries)
extern int B2R_PERIODIC;                     // _PERIODIC_
extern int B2R_ENV_3D;                       // _ENV_3D_
extern int B2R_ELEM_SIZE;                    // sizeof(CELL_DT): 4, or 1 for gol-3D.cu:6
// heat_3D_gpu_main: 0 runs the reference's active kernel, which ignores ef
// (fdm_heat_diffuse_delta{D}, heat-3D.cu:602-617); 1 runs the edge-aware
// semantics of fdm_heat_diffuse_delta1_verify (heat-3D.cu:57-158, the call
// commented out at :603) for every D: ef freezes the global faces.
extern int B2R_HEAT_VERIFY;

b2r_ctx* b2r_host_ctx();  // the context behind the globals (created by b2r_sync_init)
// b2r_sync_init's context again with another per-step radius (the elastic
// model needs 4: two radius-2 stages per step)
void b2r_host_set_radius(int radius);

// b2r_halo_sync.cu:140-183, 74-111, 121-128, 442-508, 389-431
void b2r_sync_flag(int rank);
void b2r_sync_init();
void b2r_sync_finalize();
void boundary_updating(int time_step, CELL_DT* d_grid);
void boundary_updating_direct(int time_step, CELL_DT* d_grid);

// device-resident field of all blocks (the reference's d_vx.. are per-rank)
CELL_DT* b2r_field_new();

// heat-3D.cu:583-642: R steps of delta B2R_D on one padded host block h_in of
// Ngx*Ngy*Ngz cells, then the owned z slab [R*D, Ngz-R*D) copied back;
// ef[3] in {-1,0,1} marks the block's global faces (used with B2R_HEAT_VERIFY).
double heat_3D_gpu_main(CELL_DT* h_in, int b2r_R, char* ef, int Ngx, int Ngy, int Ngz);
// gol-3D.cu:119-155: R generations of one padded host block of u8 cells, the
// whole block copied back; ef is accepted and unused, as there
double gol_3D_gpu_main(unsigned char* h_in, int b2r_R, char* ef, int Ngx, int Ngy, int Ngz);
// heat-3D.cpp:136-184: the same R steps on the host cores (b2r_cpu_heat.cpp)
float heat_3D_cpu_main(CELL_DT* h_in, int b2r_R, int b2r_D, int Ngx, int Ngy, int Ngz);
// R inner steps of the resident heat field (the device half of heat_3D_gpu_main)
double heat_3D_gpu_steps(CELL_DT** d_cur, CELL_DT** d_nxt, int b2r_R);

// eqwp.cu:515-582, 592-605, 661-714 over the elastic ABI; eqwp_field(k) is
// d_vx .. d_sigma_zz (k 0..8 in the order vx vy vz sxx syy szz sxy sxz syz)
void eqwp_cuda_init(int Ngx, int Ngy, int Ngz);
void eqwp_cuda_finalize();
double eqwp_gpu_main(int ts, int b2r_R, int Ngx, int Ngy, int Ngz, bool op_flag);
CELL_DT* eqwp_field(int which);
b2r_el_shot* eqwp_shot();

// on-disk outputs (eqwp.cu:616-650, heat-3D.hpp:105-127), written to the
// working directory.  B2R_OUTPUT selects what the drivers write: bit 0
// displacement-<ts>.bin, bit 1 the text forms (displacement-<ts>.txt for eqwp,
// one R-<R>-rank-<b>-i-<step>.txt per block for heat); eqwp_gpu_main writes the
// displacement when op_flag is set.  The reference writes one rank's padded
// block; one process here holds every block, so the displacement files carry
// the global owned grid (header ENV_DIM_X, ENV_DIM_Y, ENV_DIM_Z), and
// print_block_data_step takes the padded block to print (the reference reads
// its rank's global grid).
extern int B2R_OUTPUT;
void write_block_data_bin(int time_step, CELL_DT* p_data, int Ngx, int Ngy, int Ngz);
void write_block_data(int time_step, CELL_DT* p_data, int Ngx, int Ngy, int Ngz);
void print_block_data_step(int rank, float step, const CELL_DT* block);
