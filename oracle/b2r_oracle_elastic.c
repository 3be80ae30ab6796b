/*
 * b2r_oracle_elastic.c -- CPU oracle for the elastic FD4 velocity-stress path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the bench's
 * CPU baseline as the checker; the product never links or loads it.
 *
 * Restates eqwp.cu: the collocated 4th-order first derivative
 * (-1/12 (f[+2]-f[-2]) + 2/3 (f[+1]-f[-1]))/h (eqwp.cu:22-23, 131-133,
 * 450-458), the stress update sigma += dt*C:grad(v) with the isotropic law of
 * eqwp.cu:460-466 (lambda + 2 mu on the diagonal, lambda off it, mu for the
 * shear terms), then the velocity update v += dt*b*div(sigma) with the new
 * stresses (update order eqwp.cu:679-684), generalised to per-cell
 * lambda, mu and b = 1/rho.  Point forces are added after the velocity update
 * (v += dt*b*f) and receivers sample a velocity component after each step.
 *
 * Grids carry a zero frame of width 2 (the stencil radius of one stage):
 * (gx+4)(gy+4)(gz+4) cells, x fastest.  The state is 9 arrays in the order
 * vx, vy, vz, sxx, syy, szz, sxy, sxz, syz.
 *
 * FP32 functions evaluate exactly the GPU kernels' expressions (compiled with
 * -ffp-contract=off); FP64 functions are the error yardstick; the FP64
 * adjoint (exact transpose of the FP64 step) backs the gradient checks.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define VX 0
#define VY 1
#define VZ 2
#define SXX 3
#define SYY 4
#define SZZ 5
#define SXY 6
#define SXZ 7
#define SYZ 8

typedef struct {
    int gx, gy, gz;
    long long nx, ny, nz, n;
    long long s[3]; /* element strides of x, y, z */
} eg_t;

static eg_t mk(int gx, int gy, int gz)
{
    eg_t g;
    g.gx = gx; g.gy = gy; g.gz = gz;
    g.nx = gx + 4; g.ny = gy + 4; g.nz = gz + 4;
    g.n = g.nx * g.ny * g.nz;
    g.s[0] = 1; g.s[1] = g.nx; g.s[2] = g.nx * g.ny;
    return g;
}

static inline long long at(const eg_t* g, int x, int y, int z)
{
    return ((long long)(z + 2) * g->ny + (y + 2)) * g->nx + (x + 2);
}

/* ---------------------------------------------------------------- FP32 */
#define C2F (-1.f / 12.f)
#define C1F (2.f / 3.f)

static inline float d32(const float* f, long long c, long long s, float ih)
{
    return (C2F * (f[c + 2 * s] - f[c - 2 * s]) + C1F * (f[c + s] - f[c - s])) * ih;
}

void orc_el_stress_f32(float* const* st, const float* lam, const float* mu, int gx, int gy,
                       int gz, const float* ih, float dt)
{
    eg_t g = mk(gx, gy, gz);
    const float *vx = st[VX], *vy = st[VY], *vz = st[VZ];
#pragma omp parallel for schedule(static)
    for (int z = 0; z < gz; ++z)
        for (int y = 0; y < gy; ++y)
            for (int x = 0; x < gx; ++x) {
                long long c = at(&g, x, y, z);
                float exx = d32(vx, c, g.s[0], ih[0]);
                float eyy = d32(vy, c, g.s[1], ih[1]);
                float ezz = d32(vz, c, g.s[2], ih[2]);
                float dxy = d32(vx, c, g.s[1], ih[1]) + d32(vy, c, g.s[0], ih[0]);
                float dxz = d32(vx, c, g.s[2], ih[2]) + d32(vz, c, g.s[0], ih[0]);
                float dyz = d32(vy, c, g.s[2], ih[2]) + d32(vz, c, g.s[1], ih[1]);
                float l = lam[c], m = mu[c];
                float l2m = l + 2.0f * m;
                st[SXX][c] = st[SXX][c] + dt * (l2m * exx + l * (eyy + ezz));
                st[SYY][c] = st[SYY][c] + dt * (l2m * eyy + l * (exx + ezz));
                st[SZZ][c] = st[SZZ][c] + dt * (l2m * ezz + l * (exx + eyy));
                st[SXY][c] = st[SXY][c] + dt * (m * dxy);
                st[SXZ][c] = st[SXZ][c] + dt * (m * dxz);
                st[SYZ][c] = st[SYZ][c] + dt * (m * dyz);
            }
}

void orc_el_velocity_f32(float* const* st, const float* b, float* disp, int gx, int gy, int gz,
                         const float* ih, float dt)
{
    eg_t g = mk(gx, gy, gz);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < gz; ++z)
        for (int y = 0; y < gy; ++y)
            for (int x = 0; x < gx; ++x) {
                long long c = at(&g, x, y, z);
                float tb = dt * b[c];
                float ax = (d32(st[SXX], c, g.s[0], ih[0]) + d32(st[SXY], c, g.s[1], ih[1])) +
                           d32(st[SXZ], c, g.s[2], ih[2]);
                float ay = (d32(st[SXY], c, g.s[0], ih[0]) + d32(st[SYY], c, g.s[1], ih[1])) +
                           d32(st[SYZ], c, g.s[2], ih[2]);
                float az = (d32(st[SXZ], c, g.s[0], ih[0]) + d32(st[SYZ], c, g.s[1], ih[1])) +
                           d32(st[SZZ], c, g.s[2], ih[2]);
                st[VX][c] = st[VX][c] + tb * ax;
                st[VY][c] = st[VY][c] + tb * ay;
                st[VZ][c] = st[VZ][c] + tb * az;
                if (disp) disp[c] = disp[c] + (dt * tb) * ay;
            }
}

/* nsteps steps; sources: nsrc points (global x,y,z in src_xyz[3*i], component
 * src_comp[i]) with per-step values wav[t*nsrc + i]; receivers: nrec points,
 * component rec_comp, rec[t*nrec + r] sampled after step t. */
void orc_el_forward_f32(float* const* st, const float* lam, const float* mu, const float* b,
                        float* disp, int gx, int gy, int gz, const float* ih, float dt,
                        int nsteps, int nsrc, const int* src_xyz, const int* src_comp,
                        const float* wav, int nrec, const int* rec_xyz, int rec_comp, float* rec)
{
    eg_t g = mk(gx, gy, gz);
    for (int t = 0; t < nsteps; ++t) {
        orc_el_stress_f32(st, lam, mu, gx, gy, gz, ih, dt);
        orc_el_velocity_f32(st, b, disp, gx, gy, gz, ih, dt);
        for (int i = 0; i < nsrc; ++i) {
            long long c = at(&g, src_xyz[3 * i], src_xyz[3 * i + 1], src_xyz[3 * i + 2]);
            float tb = dt * b[c];
            float w = wav[(long long)t * nsrc + i];
            st[src_comp[i]][c] = st[src_comp[i]][c] + tb * w;
            /* the y force enters the displacement with the divergence,
             * Dt*Dt*inv_rho*(div sigma + f.y) (eqwp.cu:229-235) */
            if (disp && src_comp[i] == VY) disp[c] = disp[c] + (dt * tb) * w;
        }
        for (int r = 0; r < nrec; ++r) {
            long long c = at(&g, rec_xyz[3 * r], rec_xyz[3 * r + 1], rec_xyz[3 * r + 2]);
            rec[(long long)t * nrec + r] = st[rec_comp][c];
        }
    }
}

/* ---------------------------------------------------------------- FP64 */
#define C2D (-1.0 / 12.0)
#define C1D (2.0 / 3.0)

static inline double d64(const double* f, long long c, long long s, double ih)
{
    return (C2D * (f[c + 2 * s] - f[c - 2 * s]) + C1D * (f[c + s] - f[c - s])) * ih;
}

static void stress64(double* const* st, const double* lam, const double* mu, const eg_t* g,
                     const double* ih, double dt)
{
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g->gz; ++z)
        for (int y = 0; y < g->gy; ++y)
            for (int x = 0; x < g->gx; ++x) {
                long long c = at(g, x, y, z);
                double exx = d64(st[VX], c, g->s[0], ih[0]);
                double eyy = d64(st[VY], c, g->s[1], ih[1]);
                double ezz = d64(st[VZ], c, g->s[2], ih[2]);
                double dxy = d64(st[VX], c, g->s[1], ih[1]) + d64(st[VY], c, g->s[0], ih[0]);
                double dxz = d64(st[VX], c, g->s[2], ih[2]) + d64(st[VZ], c, g->s[0], ih[0]);
                double dyz = d64(st[VY], c, g->s[2], ih[2]) + d64(st[VZ], c, g->s[1], ih[1]);
                double l = lam[c], m = mu[c], l2m = l + 2 * m;
                st[SXX][c] += dt * (l2m * exx + l * (eyy + ezz));
                st[SYY][c] += dt * (l2m * eyy + l * (exx + ezz));
                st[SZZ][c] += dt * (l2m * ezz + l * (exx + eyy));
                st[SXY][c] += dt * m * dxy;
                st[SXZ][c] += dt * m * dxz;
                st[SYZ][c] += dt * m * dyz;
            }
}

static void velocity64(double* const* st, const double* b, const eg_t* g, const double* ih,
                       double dt)
{
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g->gz; ++z)
        for (int y = 0; y < g->gy; ++y)
            for (int x = 0; x < g->gx; ++x) {
                long long c = at(g, x, y, z);
                double tb = dt * b[c];
                st[VX][c] += tb * (d64(st[SXX], c, g->s[0], ih[0]) + d64(st[SXY], c, g->s[1], ih[1]) +
                                   d64(st[SXZ], c, g->s[2], ih[2]));
                st[VY][c] += tb * (d64(st[SXY], c, g->s[0], ih[0]) + d64(st[SYY], c, g->s[1], ih[1]) +
                                   d64(st[SYZ], c, g->s[2], ih[2]));
                st[VZ][c] += tb * (d64(st[SXZ], c, g->s[0], ih[0]) + d64(st[SYZ], c, g->s[1], ih[1]) +
                                   d64(st[SZZ], c, g->s[2], ih[2]));
            }
}

static void inject64(double* const* st, const double* b, const eg_t* g, double dt, int t,
                     int nsrc, const int* src_xyz, const int* src_comp, const double* wav)
{
    for (int i = 0; i < nsrc; ++i) {
        long long c = at(g, src_xyz[3 * i], src_xyz[3 * i + 1], src_xyz[3 * i + 2]);
        st[src_comp[i]][c] += dt * b[c] * wav[(long long)t * nsrc + i];
    }
}

void orc_el_forward_f64(double* const* st, const double* lam, const double* mu, const double* b,
                        int gx, int gy, int gz, const double* ih, double dt, int nsteps, int nsrc,
                        const int* src_xyz, const int* src_comp, const double* wav, int nrec,
                        const int* rec_xyz, int rec_comp, double* rec)
{
    eg_t g = mk(gx, gy, gz);
    for (int t = 0; t < nsteps; ++t) {
        stress64(st, lam, mu, &g, ih, dt);
        velocity64(st, b, &g, ih, dt);
        inject64(st, b, &g, dt, t, nsrc, src_xyz, src_comp, wav);
        for (int r = 0; r < nrec; ++r)
            rec[(long long)t * nrec + r] =
                st[rec_comp][at(&g, rec_xyz[3 * r], rec_xyz[3 * r + 1], rec_xyz[3 * r + 2])];
    }
}

/* Energy-like norm: sum of squares of all 9 fields over the owned cells. */
double orc_el_sumsq_f64(double* const* st, int gx, int gy, int gz)
{
    eg_t g = mk(gx, gy, gz);
    double s = 0;
    for (int k = 0; k < 9; ++k)
        for (int z = 0; z < gz; ++z)
            for (int y = 0; y < gy; ++y)
                for (int x = 0; x < gx; ++x) {
                    double v = st[k][at(&g, x, y, z)];
                    s += v * v;
                }
    return s;
}

/* ---------------------------------------------------------------- FP64 adjoint
 * Forward step t maps (v, s) -> (v', s'):  s' = s + dt C(E v);  v' = v + dt b (D s') + dt b f_t;
 * receivers d_t = v'_comp at rec cells.  J = 1/2 sum_t sum_r (d_t,r - obs_t,r)^2.
 * Reverse sweep with the exact transposes (D_a^T = -D_a on the zero-framed grid):
 *   vh += R^T res_t;  g_b += dt vh . (D s' + f_t e_src);  sh += -dt D^T-form of (b vh);
 *   g_lam, g_mu from sh and E v (pre-step v);  vh += -dt E^T (C sh).
 * Forward states are recomputed by storing all of them (small grids only).
 * Gradients are per owned cell, arrays of the framed size. */
static void zero(double* a, long long n) { memset(a, 0, sizeof(double) * n); }

void orc_el_adjoint_f64(const double* lam, const double* mu, const double* b, int gx, int gy,
                        int gz, const double* ih, double dt, int nsteps, int nsrc,
                        const int* src_xyz, const int* src_comp, const double* wav, int nrec,
                        const int* rec_xyz, int rec_comp, const double* res, double* g_lam,
                        double* g_mu, double* g_b)
{
    eg_t g = mk(gx, gy, gz);
    long long n = g.n;
    /* states[t] = state before step t (t = 0..nsteps), plus s' per step */
    double* hist = (double*)calloc((size_t)(nsteps + 1) * 9 * n, sizeof(double));
    double* sp = (double*)calloc((size_t)nsteps * 6 * n, sizeof(double));
    double* cur[9];
    for (int k = 0; k < 9; ++k) cur[k] = hist + (size_t)k * n;
    for (int t = 0; t < nsteps; ++t) {
        double* nx[9];
        for (int k = 0; k < 9; ++k) {
            nx[k] = hist + ((size_t)(t + 1) * 9 + k) * n;
            memcpy(nx[k], cur[k], sizeof(double) * n);
        }
        stress64(nx, lam, mu, &g, ih, dt);
        for (int k = 0; k < 6; ++k) memcpy(sp + ((size_t)t * 6 + k) * n, nx[3 + k], sizeof(double) * n);
        velocity64(nx, b, &g, ih, dt);
        inject64(nx, b, &g, dt, t, nsrc, src_xyz, src_comp, wav);
        for (int k = 0; k < 9; ++k) cur[k] = nx[k];
    }
    double* adj = (double*)calloc((size_t)9 * n, sizeof(double));
    double* w = (double*)calloc((size_t)9 * n, sizeof(double));
    double* vh[3] = {adj, adj + n, adj + 2 * n};
    double* sh[6] = {adj + 3 * n, adj + 4 * n, adj + 5 * n, adj + 6 * n, adj + 7 * n, adj + 8 * n};
    zero(g_lam, n); zero(g_mu, n); zero(g_b, n);
    for (int t = nsteps - 1; t >= 0; --t) {
        const double* vpre[3];
        for (int k = 0; k < 3; ++k) vpre[k] = hist + ((size_t)t * 9 + k) * n;
        const double* s1[6];
        for (int k = 0; k < 6; ++k) s1[k] = sp + ((size_t)t * 6 + k) * n;
        /* receivers at step t */
        for (int r = 0; r < nrec; ++r)
            vh[rec_comp][at(&g, rec_xyz[3 * r], rec_xyz[3 * r + 1], rec_xyz[3 * r + 2])] +=
                res[(long long)t * nrec + r];
        /* g_b: v' = v + dt b (D s') + dt b f */
        for (int z = 0; z < gz; ++z)
            for (int y = 0; y < gy; ++y)
                for (int x = 0; x < gx; ++x) {
                    long long c = at(&g, x, y, z);
                    double ax = d64(s1[0], c, g.s[0], ih[0]) + d64(s1[3], c, g.s[1], ih[1]) + d64(s1[4], c, g.s[2], ih[2]);
                    double ay = d64(s1[3], c, g.s[0], ih[0]) + d64(s1[1], c, g.s[1], ih[1]) + d64(s1[5], c, g.s[2], ih[2]);
                    double az = d64(s1[4], c, g.s[0], ih[0]) + d64(s1[5], c, g.s[1], ih[1]) + d64(s1[2], c, g.s[2], ih[2]);
                    g_b[c] += dt * (vh[0][c] * ax + vh[1][c] * ay + vh[2][c] * az);
                }
        for (int i = 0; i < nsrc; ++i) {
            long long c = at(&g, src_xyz[3 * i], src_xyz[3 * i + 1], src_xyz[3 * i + 2]);
            g_b[c] += dt * vh[src_comp[i]][c] * wav[(long long)t * nsrc + i];
        }
        /* sh += dt D^T (b vh): w_i = b vh_i; sh_ij += -dt (D_j w_i + D_i w_j) (i != j), sh_ii += -dt D_i w_i */
        for (int k = 0; k < 3; ++k)
            for (long long c = 0; c < n; ++c) w[k * n + c] = b[c] * vh[k][c];
        for (int z = 0; z < gz; ++z)
            for (int y = 0; y < gy; ++y)
                for (int x = 0; x < gx; ++x) {
                    long long c = at(&g, x, y, z);
                    const double *wx = w, *wy = w + n, *wz = w + 2 * n;
                    sh[0][c] -= dt * d64(wx, c, g.s[0], ih[0]);
                    sh[1][c] -= dt * d64(wy, c, g.s[1], ih[1]);
                    sh[2][c] -= dt * d64(wz, c, g.s[2], ih[2]);
                    sh[3][c] -= dt * (d64(wx, c, g.s[1], ih[1]) + d64(wy, c, g.s[0], ih[0]));
                    sh[4][c] -= dt * (d64(wx, c, g.s[2], ih[2]) + d64(wz, c, g.s[0], ih[0]));
                    sh[5][c] -= dt * (d64(wy, c, g.s[2], ih[2]) + d64(wz, c, g.s[1], ih[1]));
                }
        /* gradients of the stress step: s' = s + dt C(E v), v = pre-step velocity */
        for (int z = 0; z < gz; ++z)
            for (int y = 0; y < gy; ++y)
                for (int x = 0; x < gx; ++x) {
                    long long c = at(&g, x, y, z);
                    double exx = d64(vpre[0], c, g.s[0], ih[0]);
                    double eyy = d64(vpre[1], c, g.s[1], ih[1]);
                    double ezz = d64(vpre[2], c, g.s[2], ih[2]);
                    double dxy = d64(vpre[0], c, g.s[1], ih[1]) + d64(vpre[1], c, g.s[0], ih[0]);
                    double dxz = d64(vpre[0], c, g.s[2], ih[2]) + d64(vpre[2], c, g.s[0], ih[0]);
                    double dyz = d64(vpre[1], c, g.s[2], ih[2]) + d64(vpre[2], c, g.s[1], ih[1]);
                    double tr = exx + eyy + ezz;
                    double strace = sh[0][c] + sh[1][c] + sh[2][c];
                    g_lam[c] += dt * strace * tr;
                    g_mu[c] += dt * (2 * (sh[0][c] * exx + sh[1][c] * eyy + sh[2][c] * ezz) +
                                     sh[3][c] * dxy + sh[4][c] * dxz + sh[5][c] * dyz);
                }
        /* vh += dt E^T (C sh) = -dt div-form of p = C sh */
        for (long long c = 0; c < n; ++c) {
            double l = lam[c], m = mu[c], l2m = l + 2 * m;
            w[0 * n + c] = l2m * sh[0][c] + l * (sh[1][c] + sh[2][c]);
            w[1 * n + c] = l2m * sh[1][c] + l * (sh[0][c] + sh[2][c]);
            w[2 * n + c] = l2m * sh[2][c] + l * (sh[0][c] + sh[1][c]);
            w[3 * n + c] = m * sh[3][c];
            w[4 * n + c] = m * sh[4][c];
            w[5 * n + c] = m * sh[5][c];
        }
        for (int z = 0; z < gz; ++z)
            for (int y = 0; y < gy; ++y)
                for (int x = 0; x < gx; ++x) {
                    long long c = at(&g, x, y, z);
                    const double *pxx = w, *pyy = w + n, *pzz = w + 2 * n, *pxy = w + 3 * n,
                                 *pxz = w + 4 * n, *pyz = w + 5 * n;
                    vh[0][c] -= dt * (d64(pxx, c, g.s[0], ih[0]) + d64(pxy, c, g.s[1], ih[1]) + d64(pxz, c, g.s[2], ih[2]));
                    vh[1][c] -= dt * (d64(pxy, c, g.s[0], ih[0]) + d64(pyy, c, g.s[1], ih[1]) + d64(pyz, c, g.s[2], ih[2]));
                    vh[2][c] -= dt * (d64(pxz, c, g.s[0], ih[0]) + d64(pyz, c, g.s[1], ih[1]) + d64(pzz, c, g.s[2], ih[2]));
                }
        /* frame cells of the adjoint state stay zero (the frame is not a state) */
        for (int k = 0; k < 9; ++k)
            for (int z = -2; z < gz + 2; ++z)
                for (int y = -2; y < gy + 2; ++y)
                    for (int x = -2; x < gx + 2; ++x)
                        if (x < 0 || y < 0 || z < 0 || x >= gx || y >= gy || z >= gz)
                            adj[(size_t)k * n + at(&g, x, y, z)] = 0;
    }
    free(hist);
    free(sp);
    free(adj);
    free(w);
}
