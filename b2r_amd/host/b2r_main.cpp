// b2r_main.cpp -- the B2R driver of b2r.cpp:125-250 on one GPU.
//
//   b2r_main R Bx By Bz [--env X Y Z] [--D d] [--iters N] [--periodic]
//
// Checks the arguments like b2r.cpp:132-167, then runs the time loop of
// b2r.cpp:179-221 for the heat model: every R steps boundary_updating of the
// field, then R inner steps, and reports the timing lines of b2r.cpp:230-237.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "b2r_host.h"

int main(int argc, char* argv[])
{
    if (argc < 5) {
        printf("not sufficient parameter given for execution.\n");
        printf("please set parameter B2R_B_R, BLOCK_DIM_(X,Y,Z)\n");
        return 1;
    }
    B2R_R = atoi(argv[1]);
    BLOCK_DIM_X = atoi(argv[2]);
    BLOCK_DIM_Y = atoi(argv[3]);
    BLOCK_DIM_Z = atoi(argv[4]);
    int n_iter = 600;  // N_ITER (b2r.h:12)
    for (int i = 5; i < argc; ++i) {
        if (!strcmp(argv[i], "--env") && i + 3 < argc) {
            ENV_DIM_X = atoi(argv[++i]);
            ENV_DIM_Y = atoi(argv[++i]);
            ENV_DIM_Z = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--D") && i + 1 < argc) {
            B2R_D = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--iters") && i + 1 < argc) {
            n_iter = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--periodic")) {
            B2R_PERIODIC = 1;
        }
    }
    printf("arguments received, where: B2R_R = %d, BLOCK_DIM_X = %d, BLOCK_DIM_Y = %d, "
           "BLOCK_DIM_Z = %d\n", B2R_R, BLOCK_DIM_X, BLOCK_DIM_Y, BLOCK_DIM_Z);
    if (ENV_DIM_X % BLOCK_DIM_X || ENV_DIM_Y % BLOCK_DIM_Y || ENV_DIM_Z % BLOCK_DIM_Z) {
        printf("GBL_ENV_(X,Y,Z) should be perfectly divisible by the given block size "
               "BLOCK_DIM_(X,Y,Z) separately!\n");
        return 1;
    }
    b2r_sync_init();  // validates R*D <= min(B)/2 and exits like b2r.cpp:54-59
    for (int r = 0; r < BLOCK_DIM_X * BLOCK_DIM_Y * BLOCK_DIM_Z; ++r) b2r_sync_flag(r);

    CELL_DT* cur = b2r_field_new();
    CELL_DT* nxt = b2r_field_new();
    if (b2r_field_init_uniform(b2r_host_ctx(), cur, 0) < 0 ||
        b2r_field_copy(b2r_host_ctx(), nxt, cur) < 0) {
        fprintf(stderr, "B2R Error: init: %s\n", b2r_last_error());
        return -1;
    }
    b2r_synchronize(b2r_host_ctx());
    double sync_time = 0, compu_time = 0;
    auto t_all = std::chrono::steady_clock::now();
    for (int i = 0; i < n_iter; i += B2R_R) {
        auto s0 = std::chrono::steady_clock::now();
        boundary_updating(i, cur);
        b2r_synchronize(b2r_host_ctx());
        sync_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - s0).count();
        int r = n_iter - i < B2R_R ? n_iter - i : B2R_R;
        compu_time += heat_3D_gpu_steps(&cur, &nxt, r);
    }
    double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_all).count();
    printf("Execution/Model Parameters: B2R_R = %d, B2R_D: %d, ENV: %d X %d X %d, BLOCK_DIM_X = "
           "%d, BLOCK_DIM_Y = %d, BLOCK_DIM_Z = %d\n", B2R_R, B2R_D, ENV_DIM_X, ENV_DIM_Y,
           ENV_DIM_Z, BLOCK_DIM_X, BLOCK_DIM_Y, BLOCK_DIM_Z);
    printf("%d Iterations with B2R_R=%d took: %g seconds\n", n_iter, B2R_R, total);
    printf("device computation took: %g seconds\n", compu_time);
    printf("halo synchronization took: %g seconds\n", sync_time);
    printf("GLUP/s: %g\n", (double)ENV_DIM_X * ENV_DIM_Y * ENV_DIM_Z * n_iter / total / 1e9);
    b2r_sync_finalize();
    return 0;
}
