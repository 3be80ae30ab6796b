// b2r_main.cpp -- the B2R driver of b2r.cpp:125-250 on one GPU.
//
//   b2r_main R Bx By Bz [--model heat|gol|eqwp] [--env X Y Z] [--D d]
//            [--iters N] [--periodic] [--dump FILE] [--output bin|txt|both]
//
// Checks the arguments like b2r.cpp:132-167, then runs the time loop of
// b2r.cpp:179-221: every R steps boundary_updating of the model's fields
// (the nine eqwp fields, b2r.cpp:197-206), then R model steps, and reports the
// timing lines of b2r.cpp:230-237.  All Bx*By*Bz blocks live on the one GPU.
// --dump writes the final global owned grid (heat/gol: the model field; eqwp:
// v_y), x fastest, raw cells, for checking the driver end to end; the
// reference instead redirects its timing report to a file (b2r.cpp:82-97).
// --output writes the reference's files into the working directory: eqwp
// displacement-<ts>.bin / .txt after the last step group (eqwp.cu:689-709),
// heat R-<R>-rank-<b>-i-<iters>.txt per block at the end (heat-3D.hpp:105-127).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "b2r_host.h"

typedef std::chrono::steady_clock clk;

// b2r.cpp:47-67: the block geometry from the environment and the block
// counts, then b2r_sync_init (which checks 1 <= R*D <= min(B)/2 and exits
// with the library's message).  A 2D environment (ENV_DIM_Z = 1, _ENV_3D_ 0)
// has one block along z and B2R_BLOCK_SIZE_Z = 1 (b2r.cpp:63-67, 142-146).
static void simu_env_init()
{
    if (ENV_DIM_Z == 1) {
        B2R_ENV_3D = 0;
        BLOCK_DIM_Z = 1;
    }
    B2R_B_X = ENV_DIM_X / BLOCK_DIM_X;
    B2R_B_Y = ENV_DIM_Y / BLOCK_DIM_Y;
    B2R_B_Z = ENV_DIM_Z / BLOCK_DIM_Z;
    B2R_BLOCK_SIZE_X = B2R_B_X + 2 * B2R_D * B2R_R;
    B2R_BLOCK_SIZE_Y = B2R_B_Y + 2 * B2R_D * B2R_R;
    B2R_BLOCK_SIZE_Z = B2R_ENV_3D ? B2R_B_Z + 2 * B2R_D * B2R_R : 1;
    b2r_sync_init();
}

static double since(clk::time_point t0)
{
    return std::chrono::duration<double>(clk::now() - t0).count();
}

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
    std::string model = "heat";
    const char* dump = nullptr;
    for (int i = 5; i < argc; ++i) {
        if (!strcmp(argv[i], "--env") && i + 3 < argc) {
            ENV_DIM_X = atoi(argv[++i]);
            ENV_DIM_Y = atoi(argv[++i]);
            ENV_DIM_Z = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--D") && i + 1 < argc) {
            B2R_D = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--iters") && i + 1 < argc) {
            n_iter = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--model") && i + 1 < argc) {
            model = argv[++i];
        } else if (!strcmp(argv[i], "--periodic")) {
            B2R_PERIODIC = 1;
        } else if (!strcmp(argv[i], "--dump") && i + 1 < argc) {
            dump = argv[++i];
        } else if (!strcmp(argv[i], "--output") && i + 1 < argc) {
            const char* o = argv[++i];
            B2R_OUTPUT = !strcmp(o, "bin") ? 1 : !strcmp(o, "txt") ? 2 : !strcmp(o, "both") ? 3 : -1;
            if (B2R_OUTPUT < 0) {
                printf("unknown output %s (bin, txt, both)\n", o);
                return 1;
            }
        }
    }
    if (model == "gol") {
        B2R_D = 1;
        B2R_ELEM_SIZE = 1;
    } else if (model == "eqwp") {
        B2R_D = 4;  // stress + velocity stage, radius 2 each
    } else if (model != "heat") {
        printf("unknown model %s (heat, gol, eqwp)\n", model.c_str());
        return 1;
    }
    printf("arguments received, where: B2R_R = %d, BLOCK_DIM_X = %d, BLOCK_DIM_Y = %d, "
           "BLOCK_DIM_Z = %d\n", B2R_R, BLOCK_DIM_X, BLOCK_DIM_Y, BLOCK_DIM_Z);
    if (ENV_DIM_X % BLOCK_DIM_X || ENV_DIM_Y % BLOCK_DIM_Y || ENV_DIM_Z % BLOCK_DIM_Z) {
        printf("GBL_ENV_(X,Y,Z) should be perfectly divisible by the given block size "
               "BLOCK_DIM_(X,Y,Z) separately!\n");
        return 1;
    }
    if (ENV_DIM_Z == 1) {  // the 2D environment (b2r.cpp:142-146)
        B2R_ENV_3D = 0;
        BLOCK_DIM_Z = 1;
    }
    for (int r = 0; r < BLOCK_DIM_X * BLOCK_DIM_Y * BLOCK_DIM_Z; ++r) b2r_sync_flag(r);
    simu_env_init();
    b2r_ctx* ctx = b2r_host_ctx();

    CELL_DT* cur = nullptr;
    CELL_DT* nxt = nullptr;
    if (model == "eqwp") {
        eqwp_cuda_init(B2R_BLOCK_SIZE_X, B2R_BLOCK_SIZE_Y, B2R_BLOCK_SIZE_Z);
    } else {
        cur = b2r_field_new();
        nxt = b2r_field_new();
        int rc = model == "gol" ? b2r_field_init_bernoulli(ctx, cur, 1, (uint32_t)(0.3 * 16777216))
                                : b2r_field_init_uniform(ctx, cur, 0);
        if (rc < 0 || b2r_field_copy(ctx, nxt, cur) < 0) {
            fprintf(stderr, "B2R Error: init: %s\n", b2r_last_error());
            return -1;
        }
    }
    b2r_synchronize(ctx);
    double sync_time = 0, compu_time = 0;
    auto t_all = clk::now();
    for (int i = 0; i < n_iter; i += B2R_R) {
        const int r = n_iter - i < B2R_R ? n_iter - i : B2R_R;
        auto s0 = clk::now();
        if (model == "eqwp") {
            for (int k = 0; k < 9; ++k) boundary_updating(i, eqwp_field(k));
        } else {
            boundary_updating(i, cur);
        }
        b2r_synchronize(ctx);
        sync_time += since(s0);
        if (model == "eqwp") {
            compu_time += eqwp_gpu_main(i, r, B2R_BLOCK_SIZE_X, B2R_BLOCK_SIZE_Y,
                                        B2R_BLOCK_SIZE_Z, i + B2R_R >= n_iter);
        } else if (model == "heat") {
            compu_time += heat_3D_gpu_steps(&cur, &nxt, r);
        } else {
            auto c0 = clk::now();
            for (int k = 0; k < r; ++k) {
                if (b2r_gol_step(ctx, cur, nxt, k) < 0) {
                    fprintf(stderr, "B2R Error: gol step: %s\n", b2r_last_error());
                    return -1;
                }
                CELL_DT* t = cur;
                cur = nxt;
                nxt = t;
            }
            b2r_synchronize(ctx);
            compu_time += since(c0);
        }
    }
    const double total = since(t_all);
    printf("Execution/Model Parameters: model %s, B2R_R = %d, B2R_D: %d, ENV: %d X %d X %d, "
           "BLOCK_DIM_X = %d, BLOCK_DIM_Y = %d, BLOCK_DIM_Z = %d\n", model.c_str(), B2R_R, B2R_D,
           ENV_DIM_X, ENV_DIM_Y, ENV_DIM_Z, BLOCK_DIM_X, BLOCK_DIM_Y, BLOCK_DIM_Z);
    printf("%d Iterations with B2R_R=%d took: %g seconds\n", n_iter, B2R_R, total);
    printf("device computation took: %g seconds\n", compu_time);
    printf("halo synchronization took: %g seconds\n", sync_time);
    printf("GLUP/s: %g\n", (double)ENV_DIM_X * ENV_DIM_Y * ENV_DIM_Z * n_iter / total / 1e9);
    if (dump) {
        const size_t cells = (size_t)ENV_DIM_X * ENV_DIM_Y * ENV_DIM_Z;
        std::vector<unsigned char> host(cells * B2R_ELEM_SIZE);
        const void* f = model == "eqwp" ? (const void*)eqwp_field(1) : (const void*)cur;
        FILE* fp = fopen(dump, "wb");
        if (!fp || b2r_field_download(ctx, f, host.data()) < 0 ||
            fwrite(host.data(), 1, host.size(), fp) != host.size()) {
            fprintf(stderr, "B2R Error: dump %s: %s\n", dump, fp ? b2r_last_error() : "open");
            return -1;
        }
        fclose(fp);
    }
    if (model == "heat" && (B2R_OUTPUT & 2)) {
        // the library's padded block; a 2D block prints its one plane (the
        // library keeps a zero z frame around it)
        b2r_layout L;
        b2r_get_layout(ctx, &L);
        const size_t plane = (size_t)L.S[0] * L.S[1];
        std::vector<CELL_DT> blk(plane * L.S[2]);
        for (int r = 0; r < BLOCK_DIM_X * BLOCK_DIM_Y * BLOCK_DIM_Z; ++r) {
            if (b2r_block_download(ctx, cur, r, blk.data()) < 0) {
                fprintf(stderr, "B2R Error: block download: %s\n", b2r_last_error());
                return -1;
            }
            print_block_data_step(r, (float)n_iter, blk.data() + (B2R_ENV_3D ? 0 : plane * L.halo));
        }
    }
    if (model == "eqwp") eqwp_cuda_finalize();
    b2r_sync_finalize();
    return 0;
}
