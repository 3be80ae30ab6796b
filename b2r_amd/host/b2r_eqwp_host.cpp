// b2r_eqwp_host.cpp -- eqwp_cuda_init / eqwp_gpu_main (eqwp.cu:515-582,
// 661-714) as shims over the elastic C ABI, with the reference's model: a
// constant medium lambda = 32.00, mu = 174.87, 1/rho = 1/1800 (the law of
// eqwp.cu:460-466 with the constants of :15-19), Dt = 0.5, Dx = Dy = Dz = 10
// (:9-12), and a y-force of -10000 (:561-566).
//
// The force patch and schedule follow eqwp.cu: every block (rank) builds the
// patch in its own padded coordinates, x in [Ngx/2, Ngx/2 + Ngx/10), z in
// [Ngz/2, Ngz/2 + Ngz/10), on the row iy = Ngy - INIT_Y - B2R_D*R (:135, 229),
// with Ngx.. the padded block the driver passes (B + 2*B2R_D*R); and the force
// acts while the window start ts < FRC_TIME, so every inner step of such a
// window is forced (ceil(50/R)*R steps).  The patch cells are placed at their
// global positions (the owned cells of each block; a patch cell that falls in
// a block's halo is left out, as the next exchange would overwrite it there).
//
// The elastic step is two radius-2 stages, so the library runs it with a
// halo of 4R (the reference's halo of R*B2R_D is too shallow for it); when
// b2r_sync_init made the context with another radius, eqwp_cuda_init makes it
// again with radius 4 and points d_vx .. d_sigma_zz (b2r.h:60-64) at the
// fields, which the driver hands back to boundary_updating (b2r.cpp:197-206).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "b2r_host.h"

static b2r_el_shot* g_shot = nullptr;

static void die_on(int rc, const char* what)
{
    if (rc < 0) {
        fprintf(stderr, "B2R Error: %s: %s\n", what, b2r_last_error());
        exit(-1);
    }
}

b2r_el_shot* eqwp_shot() { return g_shot; }

void eqwp_cuda_init(int Ngx, int Ngy, int Ngz)
{
    if (!b2r_host_ctx()) {
        fprintf(stderr, "B2R Error: eqwp_cuda_init before b2r_sync_init\n");
        exit(-1);
    }
    if (g_shot) eqwp_cuda_finalize();
    b2r_host_set_radius(4);
    b2r_ctx* c = b2r_host_ctx();
    const float ih[3] = {0.1f, 0.1f, 0.1f};
    die_on(b2r_el_create(c, 0.5f, ih, &g_shot), "eqwp_cuda_init");
    b2r_layout L;
    die_on(b2r_get_layout(c, &L), "layout");
    const int bd[3] = {BLOCK_DIM_X, BLOCK_DIM_Y, L.nblocks / (BLOCK_DIM_X * BLOCK_DIM_Y)};
    const int env[3] = {L.B[0] * bd[0], L.B[1] * bd[1], L.B[2] * bd[2]};
    const long long n = (long long)env[0] * env[1] * env[2];
    std::vector<float> lam(n, 32.00f), mu(n, 174.87f), b(n, (float)(1.0 / 1800.0));
    die_on(b2r_el_set_medium(g_shot, lam.data(), mu.data(), b.data()), "medium");

    // the reference's padded block and its halo (from the sizes the driver passes)
    const int Ng[3] = {Ngx, Ngy, Ngz};
    int href[3];
    for (int a = 0; a < 3; ++a) href[a] = (Ng[a] - L.B[a]) / 2;
    std::vector<int> xyz, comp;
    for (int bz = 0; bz < bd[2]; ++bz)
        for (int by = 0; by < bd[1]; ++by)
            for (int bx = 0; bx < bd[0]; ++bx) {
                const int oy = Ngy - 7 - href[1] - href[1];  // owned row of iy = Ngy-INIT_Y-h
                if (oy < 0 || oy >= L.B[1]) continue;
                for (int x = Ngx / 2; x < Ngx / 2 + Ngx / 10; ++x)
                    for (int z = Ngz / 2; z < Ngz / 2 + Ngz / 10; ++z) {
                        const int ox = x - href[0], oz = z - href[2];
                        if (ox < 0 || ox >= L.B[0] || oz < 0 || oz >= L.B[2]) continue;
                        xyz.push_back(bx * L.B[0] + ox);
                        xyz.push_back(by * L.B[1] + oy);
                        xyz.push_back(bz * L.B[2] + oz);
                        comp.push_back(1);
                    }
            }
    const int nsrc = (int)comp.size(), frc_time = 50;
    const int nt = (frc_time + B2R_R - 1) / B2R_R * B2R_R;  // windows starting before 50
    if (nsrc) {
        std::vector<float> wav((size_t)nt * nsrc, -10000.0f);
        die_on(b2r_el_set_sources(g_shot, nsrc, xyz.data(), comp.data(), nt, wav.data()),
               "sources");
    }
    die_on(b2r_el_enable_displacement(g_shot), "displacement");
    CELL_DT** glob[9] = {&d_vx, &d_vy, &d_vz, &d_sigma_xx, &d_sigma_yy, &d_sigma_zz,
                         &d_sigma_xy, &d_sigma_xz, &d_sigma_yz};
    for (int k = 0; k < 9; ++k) *glob[k] = eqwp_field(k);
}

void eqwp_cuda_finalize()
{
    if (g_shot) b2r_el_destroy(g_shot);
    g_shot = nullptr;
    d_vx = d_vy = d_vz = nullptr;
    d_sigma_xx = d_sigma_xy = d_sigma_xz = d_sigma_yy = d_sigma_yz = d_sigma_zz = nullptr;
}

double eqwp_gpu_main(int ts, int b2r_R, int Ngx, int Ngy, int Ngz, bool op_flag)
{
    (void)Ngx; (void)Ngy; (void)Ngz;
    float ms = 0.0f;
    // the reference structure: stress, vx, vy, vz per step, synchronised
    die_on(b2r_el_forward(g_shot, b2r_R, 0, &ms), "eqwp_gpu_main");
    // eqwp.cu:689-709: with op_flag, the displacement goes to disk (the
    // reference has this switched off; B2R_OUTPUT switches it on)
    if (op_flag && (B2R_OUTPUT & 3)) {
        b2r_layout L;
        die_on(b2r_get_layout(b2r_host_ctx(), &L), "layout");
        const int gx = L.B[0] * BLOCK_DIM_X, gy = L.B[1] * BLOCK_DIM_Y,
                  gz = L.B[2] * (L.nblocks / (BLOCK_DIM_X * BLOCK_DIM_Y));
        std::vector<float> h((size_t)gx * gy * gz);
        die_on(b2r_el_get_state(g_shot, -1, h.data()), "displacement download");
        if (B2R_OUTPUT & 1) write_block_data_bin(ts, h.data(), gx, gy, gz);
        if (B2R_OUTPUT & 2) write_block_data(ts, h.data(), gx, gy, gz);
    }
    return ms / 1000.0;
}

CELL_DT* eqwp_field(int which)
{
    void* f = nullptr;
    die_on(b2r_el_field(g_shot, which, &f), "eqwp field");
    return (CELL_DT*)f;
}
