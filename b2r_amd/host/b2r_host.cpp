// b2r_host.cpp -- implementation of the reference-named host shims.
#include "b2r_host.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

int B2R_B_X, B2R_B_Y, B2R_B_Z;
int BLOCK_DIM_X = 1, BLOCK_DIM_Y = 1, BLOCK_DIM_Z = 1;
int B2R_BLOCK_SIZE_X, B2R_BLOCK_SIZE_Y, B2R_BLOCK_SIZE_Z;
int B2R_R = 1;
int ENV_DIM_X = 256, ENV_DIM_Y = 256, ENV_DIM_Z = 256;
int B2R_D = 1;
int B2R_PERIODIC = 0;
int B2R_ELEM_SIZE = sizeof(CELL_DT);
char b2r_halo_update[6];

static b2r_ctx* g_ctx = nullptr;

static void die_on(int rc, const char* what)
{
    if (rc < 0) {
        fprintf(stderr, "B2R Error: %s: %s\n", what, b2r_last_error());
        exit(-1);
    }
}

b2r_ctx* b2r_host_ctx() { return g_ctx; }

void b2r_sync_flag(int rank)
{
    // block coordinates and flags exactly as b2r_halo_sync.cu:142-182, plus -1
    // for a periodic wrap (b2r.h:37)
    int bx = (rank % (BLOCK_DIM_X * BLOCK_DIM_Y)) % BLOCK_DIM_X;
    int by = (rank % (BLOCK_DIM_X * BLOCK_DIM_Y)) / BLOCK_DIM_X;
    int bz = rank / (BLOCK_DIM_X * BLOCK_DIM_Y);
    int c[3] = {bx, by, bz}, n[3] = {BLOCK_DIM_X, BLOCK_DIM_Y, BLOCK_DIM_Z};
    for (int a = 0; a < 3; ++a) {
        b2r_halo_update[2 * a] = c[a] >= 1 ? 1 : (B2R_PERIODIC ? -1 : 0);
        b2r_halo_update[2 * a + 1] = c[a] < n[a] - 1 ? 1 : (B2R_PERIODIC ? -1 : 0);
    }
}

void b2r_sync_init()
{
    // simu_env_init (b2r.cpp:47-67) + b2r_sync_init: geometry and buffers
    B2R_B_X = ENV_DIM_X / BLOCK_DIM_X;
    B2R_B_Y = ENV_DIM_Y / BLOCK_DIM_Y;
    B2R_B_Z = ENV_DIM_Z / BLOCK_DIM_Z;
    B2R_BLOCK_SIZE_X = B2R_B_X + 2 * B2R_D * B2R_R;
    B2R_BLOCK_SIZE_Y = B2R_B_Y + 2 * B2R_D * B2R_R;
    B2R_BLOCK_SIZE_Z = B2R_B_Z + 2 * B2R_D * B2R_R;
    if (g_ctx) b2r_sync_finalize();
    b2r_geom g;
    g.env_x = ENV_DIM_X;
    g.env_y = ENV_DIM_Y;
    g.env_z = ENV_DIM_Z;
    g.block_dim_x = BLOCK_DIM_X;
    g.block_dim_y = BLOCK_DIM_Y;
    g.block_dim_z = BLOCK_DIM_Z;
    g.R = B2R_R;
    g.radius = B2R_D;
    g.boundary = B2R_PERIODIC ? B2R_BOUNDARY_PERIODIC : B2R_BOUNDARY_FROZEN;
    g.elem_size = B2R_ELEM_SIZE;
    die_on(b2r_create(&g, 0, &g_ctx), "b2r_sync_init");
}

void b2r_sync_finalize()
{
    if (g_ctx) b2r_destroy(g_ctx);
    g_ctx = nullptr;
}

void boundary_updating(int time_step, CELL_DT* d_grid)
{
    die_on(b2r_boundary_updating(g_ctx, time_step, d_grid), "boundary_updating");
}

void boundary_updating_direct(int time_step, CELL_DT* d_grid)
{
    void* f = d_grid;
    die_on(b2r_boundary_updating_direct(g_ctx, time_step, &f, 1), "boundary_updating_direct");
}

CELL_DT* b2r_field_new()
{
    void* f = nullptr;
    die_on(b2r_field_alloc(g_ctx, &f), "b2r_field_alloc");
    return (CELL_DT*)f;
}

static b2r_heat_params heat_params_ref(int D)
{
    // the reference's constants Alpha=Dt=0.1, Dx=Dy=1, Dz=0.5 (heat-3D.cu:9-13)
    b2r_heat_params p;
    memset(&p, 0, sizeof(p));
    p.D = D;
    p.mode = B2R_HEAT_REF;
    p.adt = 0.1 * 0.1;
    p.h2[0] = 1.0 * 1.0;
    p.h2[1] = 1.0 * 1.0;
    p.h2[2] = 0.5 * 0.5;
    return p;
}

double heat_3D_gpu_steps(CELL_DT** d_cur, CELL_DT** d_nxt, int b2r_R)
{
    b2r_heat_params p = heat_params_ref(B2R_D);
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < b2r_R; ++i) {
        die_on(b2r_heat_step(g_ctx, *d_cur, *d_nxt, &p, i), "heat step");
        CELL_DT* t = *d_cur;
        *d_cur = *d_nxt;
        *d_nxt = t;
    }
    die_on(b2r_synchronize(g_ctx), "heat sync");
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// One padded host block: a private single-block context whose frozen faces
// come from ef, R inner steps, copy back (heat-3D.cu:583-642).
template <typename T>
static double block_main(T* h_in, int b2r_R, const char* ef, int Ngx, int Ngy, int Ngz, int D,
                         bool heat)
{
    int h = D * b2r_R;
    b2r_geom g;
    g.env_x = Ngx - 2 * h;
    g.env_y = Ngy - 2 * h;
    g.env_z = Ngz - 2 * h;
    g.block_dim_x = g.block_dim_y = g.block_dim_z = 1;
    g.R = b2r_R;
    g.radius = D;
    g.boundary = B2R_BOUNDARY_FROZEN;
    g.elem_size = sizeof(T);
    b2r_ctx* c = nullptr;
    die_on(b2r_create(&g, 0, &c), "block context");
    int mask = 0;
    for (int a = 0; a < 3; ++a) {
        if (ef && ef[a] == -1) mask |= 1 << (2 * a);
        if (ef && ef[a] == 1) mask |= 1 << (2 * a + 1);
    }
    die_on(b2r_set_frozen_faces(c, mask), "frozen faces");
    void *a = nullptr, *b = nullptr;
    die_on(b2r_field_alloc(c, &a), "alloc");
    die_on(b2r_field_alloc(c, &b), "alloc");
    die_on(b2r_block_upload(c, a, 0, h_in), "upload");
    die_on(b2r_field_copy(c, b, a), "copy");
    b2r_heat_params p = heat_params_ref(D);
    die_on(b2r_synchronize(c), "sync");
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < b2r_R; ++i) {
        int rc = heat ? b2r_heat_step(c, a, b, &p, i) : b2r_gol_step(c, a, b, i);
        die_on(rc, "step");
        void* t = a;
        a = b;
        b = t;
    }
    die_on(b2r_synchronize(c), "sync");
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    die_on(b2r_block_download(c, a, 0, h_in), "download");
    b2r_destroy(c);
    return secs;
}

double heat_3D_gpu_main(CELL_DT* h_in, int b2r_R, char* ef, int Ngx, int Ngy, int Ngz)
{
    return block_main<CELL_DT>(h_in, b2r_R, ef, Ngx, Ngy, Ngz, B2R_D, true);
}

double gol_3D_gpu_main(unsigned char* h_in, int b2r_R, char* ef, int Ngx, int Ngy, int Ngz)
{
    return block_main<unsigned char>(h_in, b2r_R, ef, Ngx, Ngy, Ngz, 1, false);
}
