// b2r_shaped_main.cpp -- the reference's driver b2r.cpp (b2r.cpp:25-250) with
// its MPI calls taken out and nothing else changed in its use of the B2R
// interface: the globals of b2r.h:51-87 are defined here, simu_env_init is
// b2r.cpp:47-67, and the time loop hands d_vx .. d_sigma_zz to
// boundary_updating and calls eqwp_gpu_main as b2r.cpp:179-221 does.  It links
// against libb2r_host.so built with this directory's b2r_config.h, so it shows
// that a b2r.cpp-shaped program drives the library through the reference's
// names alone.  One process owns every block (rank 0 of np = the block
// count).  Test additions: N_ITER from the command line (argv[5]) and a dump
// of v_y's global owned grid (argv[6]) through the C ABI after the loop.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "b2r_config.h"
#include "b2r_hip.h"

using namespace std;

// b2r.h:51-87 (without the MPI requests)
#if _ENV_3D_
char b2r_halo_update[6];
#else
char b2r_halo_update[4];
#endif
CELL_DT *d_vx, *d_vy, *d_vz;
CELL_DT *d_sigma_xx, *d_sigma_xy, *d_sigma_xz, *d_sigma_yy, *d_sigma_yz, *d_sigma_zz;
int B2R_B_X, B2R_B_Y, B2R_B_Z;
int BLOCK_DIM_X;
int BLOCK_DIM_Y;
int BLOCK_DIM_Z;
int B2R_BLOCK_SIZE_X;
int B2R_BLOCK_SIZE_Y;
int B2R_BLOCK_SIZE_Z;
int B2R_R;
int item_size = sizeof(CELL_DT);

// b2r.cpp:30-35 and the model entry it calls (eqwp.cu:661)
extern void boundary_updating(int time_step, CELL_DT* d_grid);
extern void b2r_sync_init();
extern void b2r_sync_flag(int rank);
extern void eqwp_cuda_init(int Ngx, int Ngy, int Ngz);
extern void b2r_sync_finalize();
extern double eqwp_gpu_main(int ts, int b2r_R, int Ngx, int Ngy, int Ngz, bool op_flag);
extern void eqwp_cuda_finalize();
extern b2r_ctx* b2r_host_ctx();

void simu_env_init(int np, int rank)
{
    (void)np;
    B2R_B_X = ENV_DIM_X / BLOCK_DIM_X;
    B2R_B_Y = ENV_DIM_Y / BLOCK_DIM_Y;
    B2R_B_Z = ENV_DIM_Z / BLOCK_DIM_Z;
#if _ENV_3D_
    if (B2R_R * B2R_D > min(B2R_B_X, min(B2R_B_Y, B2R_B_Z)) / 2) {
        if (rank == 0)
            cout << "should meet condition: 1 <= B2R_R*B2R_D <= min(B2R_B_X, B2R_B_Y, B2R_B_Z)/2"
                 << endl;
        exit(1);
    }
#endif
    B2R_BLOCK_SIZE_X = (B2R_B_X + 2 * B2R_D * B2R_R);
    B2R_BLOCK_SIZE_Y = (B2R_B_Y + 2 * B2R_D * B2R_R);
#if _ENV_3D_
    B2R_BLOCK_SIZE_Z = (B2R_B_Z + 2 * B2R_D * B2R_R);
#else
    B2R_BLOCK_SIZE_Z = 1;
#endif
    b2r_sync_init();
}

int main(int argc, char* argv[])
{
    int rank = 0;
    if (argc < 7) {
        cout << "usage: b2r_dropin R Bx By Bz N_ITER dump_file" << endl;
        exit(1);
    }
    B2R_R = atoi(argv[1]);
    BLOCK_DIM_X = atoi(argv[2]);
    BLOCK_DIM_Y = atoi(argv[3]);
#if _ENV_3D_
    BLOCK_DIM_Z = atoi(argv[4]);
#else
    BLOCK_DIM_Z = 1;
#endif
    const int N_ITER = atoi(argv[5]);
    const int np = BLOCK_DIM_X * BLOCK_DIM_Y * BLOCK_DIM_Z;
    if (ENV_DIM_X % BLOCK_DIM_X != 0 || ENV_DIM_Y % BLOCK_DIM_Y != 0 || ENV_DIM_Z % BLOCK_DIM_Z != 0) {
        cout << "GBL_ENV_(X,Y,Z) should be perfectly divisible by the given block size "
                "BLOCK_DIM_(X,Y,Z) separately!" << endl;
        exit(1);
    }
    b2r_sync_flag(rank);
    simu_env_init(np, rank);
    eqwp_cuda_init(B2R_BLOCK_SIZE_X, B2R_BLOCK_SIZE_Y, B2R_BLOCK_SIZE_Z);

    double device_compu_time = 0.0;
    for (int i = 0; i < N_ITER; i += B2R_R) {
        boundary_updating(i, d_vx);
        boundary_updating(i, d_vy);
        boundary_updating(i, d_vz);

        boundary_updating(i, d_sigma_xx);
        boundary_updating(i, d_sigma_xy);
        boundary_updating(i, d_sigma_xz);
        boundary_updating(i, d_sigma_yy);
        boundary_updating(i, d_sigma_yz);
        boundary_updating(i, d_sigma_zz);

        device_compu_time += eqwp_gpu_main(i, B2R_R, B2R_BLOCK_SIZE_X, B2R_BLOCK_SIZE_Y,
                                           B2R_BLOCK_SIZE_Z, (i + B2R_R) >= (N_ITER));
    }
    cout << "Execution/Model Parameters: B2R_R = " << B2R_R << ", B2R_D: " << B2R_D << ", ENV: "
         << B2R_B_X * BLOCK_DIM_X << " X " << B2R_B_Y * BLOCK_DIM_Y << " X "
         << B2R_B_Z * BLOCK_DIM_Z << endl;
    cout << "device computation took: " << device_compu_time << " seconds" << endl;
    cout << "halo flags:";
    for (size_t k = 0; k < sizeof(b2r_halo_update); ++k) cout << " " << (int)b2r_halo_update[k];
    cout << endl;

    std::vector<float> vy((size_t)ENV_DIM_X * ENV_DIM_Y * ENV_DIM_Z);
    if (b2r_field_download(b2r_host_ctx(), d_vy, vy.data()) != 0) {
        cerr << "dump: " << b2r_last_error() << endl;
        return 2;
    }
    FILE* f = fopen(argv[6], "wb");
    if (!f || fwrite(vy.data(), 4, vy.size(), f) != vy.size()) return 3;
    fclose(f);
    eqwp_cuda_finalize();
    b2r_sync_finalize();
    return 0;
}
