// b2r_host.cpp -- implementation of the reference-named host shims.
#include "b2r_host.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

extern const int b2r_cfg_env[3], b2r_cfg_d, b2r_cfg_periodic, b2r_cfg_env_3d;

// the globals of b2r.h (a driver that includes b2r.h brings its own copies,
// which these references then bind to)
int B2R_B_X, B2R_B_Y, B2R_B_Z;
int BLOCK_DIM_X = 1, BLOCK_DIM_Y = 1, BLOCK_DIM_Z = 1;
int B2R_BLOCK_SIZE_X, B2R_BLOCK_SIZE_Y, B2R_BLOCK_SIZE_Z;
int B2R_R = 1;
char b2r_halo_update[6];
CELL_DT *d_vx, *d_vy, *d_vz;
CELL_DT *d_sigma_xx, *d_sigma_xy, *d_sigma_xz, *d_sigma_yy, *d_sigma_yz, *d_sigma_zz;

// the switches of b2r_config.h
int ENV_DIM_X = b2r_cfg_env[0], ENV_DIM_Y = b2r_cfg_env[1], ENV_DIM_Z = b2r_cfg_env[2];
int B2R_D = b2r_cfg_d;
int B2R_PERIODIC = b2r_cfg_periodic;
int B2R_ENV_3D = b2r_cfg_env_3d;
int B2R_ELEM_SIZE = sizeof(CELL_DT);
int B2R_HEAT_VERIFY = 0;

static b2r_ctx* g_ctx = nullptr;

static void die_on(int rc, const char* what)
{
    if (rc < 0) {
        fprintf(stderr, "B2R Error: %s: %s\n", what, b2r_last_error());
        exit(-1);
    }
}

static void die(const char* what)
{
    fprintf(stderr, "B2R Error: %s\n", what);
    exit(-1);
}

b2r_ctx* b2r_host_ctx() { return g_ctx; }

void b2r_sync_flag(int rank)
{
    // block coordinates and flags as b2r_halo_sync.cu:142-182, plus -1 for a
    // periodic wrap (b2r.h:37); a 2D build has 4 entries (b2r.h:38-42)
    int bx = (rank % (BLOCK_DIM_X * BLOCK_DIM_Y)) % BLOCK_DIM_X;
    int by = (rank % (BLOCK_DIM_X * BLOCK_DIM_Y)) / BLOCK_DIM_X;
    int bz = rank / (BLOCK_DIM_X * BLOCK_DIM_Y);
    int c[3] = {bx, by, bz}, n[3] = {BLOCK_DIM_X, BLOCK_DIM_Y, BLOCK_DIM_Z};
    for (int a = 0; a < (B2R_ENV_3D ? 3 : 2); ++a) {
        b2r_halo_update[2 * a] = c[a] >= 1 ? 1 : (B2R_PERIODIC ? -1 : 0);
        b2r_halo_update[2 * a + 1] = c[a] < n[a] - 1 ? 1 : (B2R_PERIODIC ? -1 : 0);
    }
}

// The library context of b2r_sync_init's geometry with a given radius.
static void create_ctx(int radius)
{
    b2r_geom g;
    g.env_x = B2R_B_X * BLOCK_DIM_X;
    g.env_y = B2R_B_Y * BLOCK_DIM_Y;
    // 2D (b2r.cpp:63-67): one plane, no z halo in the reference
    g.env_z = B2R_BLOCK_SIZE_Z == 1 ? 1 : B2R_B_Z * BLOCK_DIM_Z;
    g.block_dim_x = BLOCK_DIM_X;
    g.block_dim_y = BLOCK_DIM_Y;
    g.block_dim_z = B2R_BLOCK_SIZE_Z == 1 ? 1 : BLOCK_DIM_Z;
    g.R = B2R_R;
    g.radius = radius;
    g.boundary = B2R_PERIODIC ? B2R_BOUNDARY_PERIODIC : B2R_BOUNDARY_FROZEN;
    g.elem_size = B2R_ELEM_SIZE;
    if (g_ctx) b2r_destroy(g_ctx);
    g_ctx = nullptr;
    die_on(b2r_create(&g, 0, &g_ctx), "b2r_sync_init");
}

void b2r_sync_init()
{
    // b2r_halo_sync.cu:74-111 reads the globals simu_env_init has set
    // (b2r.cpp:49-67): B = ENV / BLOCK_DIM and S = B + 2*B2R_D*R, so the
    // per-step radius is (S - B) / (2R)
    if (B2R_B_X <= 0 || B2R_B_Y <= 0 || B2R_B_Z <= 0 || B2R_R <= 0 || BLOCK_DIM_X <= 0 ||
        BLOCK_DIM_Y <= 0 || BLOCK_DIM_Z <= 0)
        die("b2r_sync_init: B2R_B_*, BLOCK_DIM_* and B2R_R must be set first (simu_env_init, "
            "b2r.cpp:47-67)");
    const int hx = B2R_BLOCK_SIZE_X - B2R_B_X;
    if (hx <= 0 || hx % (2 * B2R_R) || B2R_BLOCK_SIZE_Y - B2R_B_Y != hx ||
        (B2R_BLOCK_SIZE_Z != 1 && B2R_BLOCK_SIZE_Z - B2R_B_Z != hx))
        die("b2r_sync_init: B2R_BLOCK_SIZE_* must be B2R_B_* + 2*B2R_D*B2R_R (b2r.cpp:61-67)");
    create_ctx(hx / (2 * B2R_R));
}

void b2r_host_set_radius(int radius)
{
    if (!g_ctx) die("b2r_sync_init has not run");
    b2r_layout L;
    die_on(b2r_get_layout(g_ctx, &L), "layout");
    if (L.halo != radius * B2R_R) create_ctx(radius);
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

// The reference's constants Alpha = Dt = 0.1, Dx = Dy = 1, Dz = 0.5
// (heat-3D.cu:9-13) in the reference expression (REF mode).  A 2D grid has no
// z term: F32 mode with k = Alpha*Dt/h^2 on x and y.
static b2r_heat_params heat_params_ref(int D, bool two_d)
{
    b2r_heat_params p;
    memset(&p, 0, sizeof(p));
    p.D = D;
    p.adt = 0.1 * 0.1;
    p.h2[0] = 1.0 * 1.0;
    p.h2[1] = 1.0 * 1.0;
    p.h2[2] = 0.5 * 0.5;
    if (two_d) {
        p.mode = B2R_HEAT_F32;
        p.k[0] = (float)(p.adt / p.h2[0]);
        p.k[1] = (float)(p.adt / p.h2[1]);
        p.k[2] = 0.0f;
    } else {
        p.mode = B2R_HEAT_REF;
    }
    return p;
}

double heat_3D_gpu_steps(CELL_DT** d_cur, CELL_DT** d_nxt, int b2r_R)
{
    b2r_layout L;
    die_on(b2r_get_layout(g_ctx, &L), "layout");
    b2r_heat_params p = heat_params_ref(B2R_D, B2R_BLOCK_SIZE_Z == 1);
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

// One padded host block on a private single-block context: R inner steps,
// then the copy back (heat-3D.cu:583-642, gol-3D.cu:119-155).  The faces are
// open (the reference's plain kernels update the whole inner-step region of
// the block, heat-3D.cu:191, gol-3D.cu:50) unless frozen_from_ef, which
// freezes the faces ef marks (-1: minus, 1: plus).  copy_z0 > 0 copies back
// only the z planes [copy_z0, Ngz - copy_z0) (heat-3D.cu:633-635).
template <typename T>
static double block_main(T* h_in, int b2r_R, const char* ef, int Ngx, int Ngy, int Ngz, int D,
                         bool heat, bool frozen_from_ef, int copy_z0)
{
    if (Ngz == 1) die("the model entries take 3D blocks (Ngz > 1)");
    const int h = D * b2r_R;
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
    for (int a = 0; a < 3 && frozen_from_ef && ef; ++a) {
        if (ef[a] == -1) mask |= 1 << (2 * a);
        if (ef[a] == 1) mask |= 1 << (2 * a + 1);
    }
    die_on(b2r_set_frozen_faces(c, mask), "frozen faces");
    void *a = nullptr, *b = nullptr;
    die_on(b2r_field_alloc(c, &a), "alloc");
    die_on(b2r_field_alloc(c, &b), "alloc");
    die_on(b2r_block_upload(c, a, 0, h_in), "upload");
    die_on(b2r_field_copy(c, b, a), "copy");
    b2r_heat_params p = heat_params_ref(D, false);
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
    const size_t plane = (size_t)Ngx * Ngy;
    if (copy_z0 > 0) {
        std::vector<T> tmp(plane * Ngz);
        die_on(b2r_block_download(c, a, 0, tmp.data()), "download");
        memcpy(h_in + plane * copy_z0, tmp.data() + plane * copy_z0,
               sizeof(T) * plane * (Ngz - 2 * copy_z0));
    } else {
        die_on(b2r_block_download(c, a, 0, h_in), "download");
    }
    b2r_destroy(c);
    return secs;
}

double heat_3D_gpu_main(CELL_DT* h_in, int b2r_R, char* ef, int Ngx, int Ngy, int Ngz)
{
    return block_main<CELL_DT>(h_in, b2r_R, ef, Ngx, Ngy, Ngz, B2R_D, true, B2R_HEAT_VERIFY != 0,
                               b2r_R * B2R_D);
}

double gol_3D_gpu_main(unsigned char* h_in, int b2r_R, char* ef, int Ngx, int Ngy, int Ngz)
{
    return block_main<unsigned char>(h_in, b2r_R, ef, Ngx, Ngy, Ngz, 1, false, false, 0);
}

// heat-3D.cpp:136-184: R inner steps of delta-b2r_D on the host cores
// (b2r_heat_cpu_step, the reference expression), ping-pong, and the owned z
// slab [R*D, Ngz-R*D) of the last result copied back into h_in when it is the
// second buffer.  The reference copies with the macro B2R_D in place of its
// b2r_D argument; this uses b2r_D.  Returns wall seconds.
float heat_3D_cpu_main(CELL_DT* h_in, int b2r_R, int b2r_D, int Ngx, int Ngy, int Ngz)
{
    auto t0 = std::chrono::steady_clock::now();
    b2r_heat_params p = heat_params_ref(b2r_D, false);
    const size_t n = (size_t)Ngx * Ngy * Ngz, plane = (size_t)Ngx * Ngy;
    std::vector<CELL_DT> other(h_in, h_in + n);
    CELL_DT* in = h_in;
    CELL_DT* out = other.data();
    for (int i = 0; i < b2r_R; ++i) {
        die_on(b2r_heat_cpu_step(&p, in, out, Ngx, Ngy, Ngz, i, 0), "heat_3D_cpu_main");
        CELL_DT* t = in;
        in = out;
        out = t;
    }
    if (in != h_in) {
        const size_t z0 = (size_t)b2r_R * b2r_D;
        memcpy(h_in + z0 * plane, in + z0 * plane, sizeof(CELL_DT) * plane * (Ngz - 2 * z0));
    }
    return (float)std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
