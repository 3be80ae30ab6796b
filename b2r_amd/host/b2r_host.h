// b2r_host.h -- the reference's host-side interface, kept as thin C++ shims over
// the C ABI (include/b2r_hip.h).  Names, arguments and error behaviour follow
// b2r.cpp, b2r.h, b2r_halo_sync.cu, heat-3D.cu and gol-3D.cu: the shims print
// the library's message and exit(-1) on failure, as cudaErrchk does
// (b2r_halo_sync.cu:22-28).  One process drives all Bx*By*Bz blocks on one GPU,
// so "rank" means block index and the fields are device-resident multi-block
// fields of the library.
#pragma once

#include "../../include/b2r_hip.h"

typedef float CELL_DT;  // b2r_config.h:17

// geometry globals (b2r.h:72-87)
extern int B2R_B_X, B2R_B_Y, B2R_B_Z;
extern int BLOCK_DIM_X, BLOCK_DIM_Y, BLOCK_DIM_Z;
extern int B2R_BLOCK_SIZE_X, B2R_BLOCK_SIZE_Y, B2R_BLOCK_SIZE_Z;
extern int B2R_R;
extern int ENV_DIM_X, ENV_DIM_Y, ENV_DIM_Z;  // compile-time in b2r_config.h:6-8
extern int B2R_D;                            // compile-time in b2r_config.h:10
extern int B2R_PERIODIC;                     // _PERIODIC_ (b2r_config.h:12)
extern int B2R_ELEM_SIZE;                    // sizeof(CELL_DT): 4, or 1 for gol-3D.cu:6
extern char b2r_halo_update[6];              // b2r.h:37-42

b2r_ctx* b2r_host_ctx();  // the context behind the globals (created by b2r_sync_init)

// b2r_halo_sync.cu:140-183, 74-111, 121-128, 442-508, 389-431
void b2r_sync_flag(int rank);
void b2r_sync_init();
void b2r_sync_finalize();
void boundary_updating(int time_step, CELL_DT* d_grid);
void boundary_updating_direct(int time_step, CELL_DT* d_grid);

// device-resident field of all blocks (the reference's d_vx.. are per-rank)
CELL_DT* b2r_field_new();

// heat-3D.cu:583-642: R steps of delta B2R_D on one padded host block h_in of
// Ngx*Ngy*Ngz cells; ef[3] in {-1,0,1} freezes the minus/plus global face.
double heat_3D_gpu_main(CELL_DT* h_in, int b2r_R, char* ef, int Ngx, int Ngy, int Ngz);
// gol-3D.cu:119-155, same contract for u8 cells
double gol_3D_gpu_main(unsigned char* h_in, int b2r_R, char* ef, int Ngx, int Ngy, int Ngz);
// R inner steps of the resident heat field (the device half of heat_3D_gpu_main)
double heat_3D_gpu_steps(CELL_DT** d_cur, CELL_DT** d_nxt, int b2r_R);

// eqwp.cu:515-582, 592-605, 661-714 over the elastic ABI (B2R_D must be 4:
// two radius-2 stages per step); eqwp_field(k) is d_vx .. d_sigma_zz (k 0..8)
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
