/*
 * b2r_oracle.c -- CPU oracle for the B2R stencil hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the C++
 * host driver, the Python host mirror) links, loads or calls this file; only the
 * tests, the smoke check and the bench's cpu_baseline leg do, as the checker.
 *
 * It restates the reference's arithmetic in plain C, compiled with
 * -ffp-contract=off so that every FP32 expression rounds exactly as written:
 *   - heat fdm_heat_diffuse_delta1..4 (heat-3D.cu:173-521, heat-3D.cpp:48-62),
 *     in FP64 (the error yardstick) and in FP32 with the evaluation order the
 *     GPU kernels use (F32 mode) or the reference's own expression with its
 *     double promotion (REF mode, heat-3D.cpp:57-60);
 *   - the reference CPU step cpu_fdm_heat_diffuse1 on one padded block, with its
 *     loop bounds and reversed z order (heat-3D.cpp:48-62);
 *   - the 3D Game of Life step with a total rule (gol-3D.cu:83-99);
 *   - the B2R halo pack / unpack and the ordered X->Y->Z exchange
 *     (b2r_halo_sync.cu:195-379, 442-508; neighbours b2r_comm.hpp:25-28,
 *     b2r_halo_sync.cu:140-183);
 *   - SplitMix64 seeded generators shared with the GPU;
 *   - an FP64 adjoint of the heat time loop (no reference counterpart).
 *
 * Pinning (see tests/): delta1 REF mode is compared bit-for-bit with the
 * reference's own heat-3D.cpp compiled under oracle/_ref; FP64 heat against the
 * closed-form Dirichlet sine eigenmode and periodic Fourier modes, polynomial
 * exactness and conservation; GoL against known patterns, a 27-cell brute force
 * and translation equivariance; the exchange against global-index maps; the
 * heat adjoint against central finite differences of the FP64 forward and the
 * dot-product identity <A^N x, y> = <x, (A^T)^N y>.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* SplitMix64 generators                                                     */
/* ------------------------------------------------------------------------- */
static inline uint64_t sm64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

uint64_t orc_splitmix64(uint64_t x) { return sm64(x); }

/* 24-bit draw for cell gid of stream seed. */
static inline uint32_t draw24(uint64_t seed, uint64_t gid)
{
    return (uint32_t)(sm64(sm64(seed) + gid) >> 40);
}

void orc_gen_uniform(float* out, long long n, uint64_t seed)
{
    for (long long i = 0; i < n; ++i)
        out[i] = (float)draw24(seed, (uint64_t)i) * (1.0f / 16777216.0f);
}

void orc_gen_bernoulli(uint8_t* out, long long n, uint64_t seed, uint32_t p_q24)
{
    for (long long i = 0; i < n; ++i)
        out[i] = draw24(seed, (uint64_t)i) < p_q24 ? 1 : 0;
}

/* Separable Gaussian blob sum, order fixed: t=((amp*ex)*ey)*ez; acc+=t. */
void orc_gen_blobs(float* out, int gx, int gy, int gz, int nb, const double* amp,
                   const double* ex, const double* ey, const double* ez)
{
    for (int z = 0; z < gz; ++z)
        for (int y = 0; y < gy; ++y)
            for (int x = 0; x < gx; ++x) {
                double acc = 0.0;
                for (int k = 0; k < nb; ++k) {
                    double t = amp[k] * ex[(long long)k * gx + x];
                    t = t * ey[(long long)k * gy + y];
                    t = t * ez[(long long)k * gz + z];
                    acc = acc + t;
                }
                out[((long long)z * gy + y) * gx + x] = (float)acc;
            }
}

/* ------------------------------------------------------------------------- */
/* Heat                                                                      */
/* ------------------------------------------------------------------------- */
/* Second-difference taps c[0..D] (c[0] the centre), heat-3D.cu:15-28 and the
 * 7-point weights of delta1.  FP32 values exactly as the reference's literals. */
static void coef_f32(int D, float c[5])
{
    memset(c, 0, 5 * sizeof(float));
    switch (D) {
    case 1: c[0] = -2.0f; c[1] = 1.0f; break;
    case 2: c[0] = -2.5f; c[1] = 4.f / 3.f; c[2] = -1.f / 12.f; break;
    case 3: c[0] = -49.f / 18.f; c[1] = 1.5f; c[2] = -3.f / 20.f; c[3] = 1.f / 90.f; break;
    case 4: c[0] = -205.f / 72.f; c[1] = 1.6f; c[2] = -0.2f; c[3] = 8.f / 315.f;
            c[4] = -1.f / 560.f; break;
    }
}

static void coef_f64(int D, double c[5])
{
    memset(c, 0, 5 * sizeof(double));
    switch (D) {
    case 1: c[0] = -2.0; c[1] = 1.0; break;
    case 2: c[0] = -5.0 / 2.0; c[1] = 4.0 / 3.0; c[2] = -1.0 / 12.0; break;
    case 3: c[0] = -49.0 / 18.0; c[1] = 3.0 / 2.0; c[2] = -3.0 / 20.0; c[3] = 1.0 / 90.0; break;
    case 4: c[0] = -205.0 / 72.0; c[1] = 8.0 / 5.0; c[2] = -1.0 / 5.0; c[3] = 8.0 / 315.0;
            c[4] = -1.0 / 560.0; break;
    }
}

void orc_heat_coef_f32(int D, float* c) { coef_f32(D, c); }
void orc_heat_coef_f64(int D, double* c) { coef_f64(D, c); }

/* Index of global cell (x,y,z) in the oracle grid.  periodic is 0 (frozen), 1
 * (every axis wraps) or a per-axis mask (bit a: axis a wraps; 3 is a 2D grid
 * of one z plane that wraps x and y).  A frozen axis carries a frame of width
 * D (n + 2D cells, coordinate in [-D, n+D)); a periodic axis has n cells and
 * wraps. */
typedef struct {
    int gx, gy, gz, D, periodic;  /* periodic: the per-axis mask */
    long long nx, ny, nz;
} grid_t;

static inline int wrap_or_frame(int v, int n, int D, int per)
{
    return per ? ((v % n) + n) % n : v + D;
}

static inline long long gidx(const grid_t* g, int x, int y, int z)
{
    const int pm = g->periodic;
    x = wrap_or_frame(x, g->gx, g->D, pm & 1);
    y = wrap_or_frame(y, g->gy, g->D, (pm >> 1) & 1);
    z = wrap_or_frame(z, g->gz, g->D, (pm >> 2) & 1);
    return ((long long)z * g->ny + y) * g->nx + x;
}

static grid_t mkgrid(int gx, int gy, int gz, int D, int periodic)
{
    grid_t g;
    const int pm = periodic == 1 ? 7 : periodic;
    g.gx = gx; g.gy = gy; g.gz = gz; g.D = D; g.periodic = pm;
    g.nx = (pm & 1) ? gx : gx + 2 * D;
    g.ny = (pm & 2) ? gy : gy + 2 * D;
    g.nz = (pm & 4) ? gz : gz + 2 * D;
    return g;
}

static long long grid_elems(const grid_t* g) { return g->nx * g->ny * g->nz; }

/* FP64 heat: u' = u + sum_a k_a * (c0 u + sum_j c_j (u[-j] + u[+j])). */
void orc_heat_f64(double* u, int gx, int gy, int gz, int D, const double* k,
                  int nsteps, int periodic)
{
    grid_t g = mkgrid(gx, gy, gz, D, periodic);
    long long n = grid_elems(&g);
    double c[5];
    coef_f64(D, c);
    double* v = (double*)malloc(sizeof(double) * n);
    memcpy(v, u, sizeof(double) * n);
    double* src = u;
    double* dst = v;
    for (int s = 0; s < nsteps; ++s) {
#pragma omp parallel for schedule(static)
        for (int z = 0; z < gz; ++z)
            for (int y = 0; y < gy; ++y)
                for (int x = 0; x < gx; ++x) {
                    double uc = src[gidx(&g, x, y, z)];
                    double lx = c[0] * uc, ly = c[0] * uc, lz = c[0] * uc;
                    for (int j = 1; j <= D; ++j) {
                        lx += c[j] * (src[gidx(&g, x - j, y, z)] + src[gidx(&g, x + j, y, z)]);
                        ly += c[j] * (src[gidx(&g, x, y - j, z)] + src[gidx(&g, x, y + j, z)]);
                        lz += c[j] * (src[gidx(&g, x, y, z - j)] + src[gidx(&g, x, y, z + j)]);
                    }
                    dst[gidx(&g, x, y, z)] = uc + k[0] * lx + k[1] * ly + k[2] * lz;
                }
        double* t = src; src = dst; dst = t;
    }
    if (src != u) memcpy(u, src, sizeof(double) * n);
    free(v);
}

/* FP32 L_a with the fixed tap order: outermost pair first, centre last. */
static inline float lap_f32(const float* c, int D, float uc, const float* m, const float* p)
{
    float acc = c[D] * (m[D] + p[D]);
    for (int j = D - 1; j >= 1; --j) acc = acc + c[j] * (m[j] + p[j]);
    return acc + c[0] * uc;
}

/* F32-mode L_a: the taps of lap_f32 with each product fused with its sum. */
static inline float lap_fma(const float* c, int D, float uc, const float* m, const float* p)
{
    float acc = c[D] * (m[D] + p[D]);
    for (int j = D - 1; j >= 1; --j) acc = fmaf(c[j], m[j] + p[j], acc);
    return fmaf(c[0], uc, acc);
}

/* Reference form of L_a (heat-3D.cpp:58-60 for D=1, heat-3D.cu:316-318 for D>1). */
static inline float lap_ref(const float* c, int D, float uc, const float* m, const float* p)
{
    if (D == 1) return (m[1] - 2.0f * uc) + p[1];
    return lap_f32(c, D, uc, m, p);
}

/* FP32 heat with exactly the GPU kernels' arithmetic.
 * mode 0 (F32): L_a = c_D*(m_D+p_D) + ... + c_1*(m_1+p_1) + c_0*u (outermost
 *               first) and t = kx*Lx + ky*Ly + kz*Lz, each product fused with
 *               its sum (fmaf, one rounding), then u' = u + t
 * mode 1 (REF): u' = (float)((double)u + adt*(((double)Lx/h2x + (double)Ly/h2y)
 *                                              + (double)Lz/h2z)) */
void orc_heat_f32(float* u, int gx, int gy, int gz, int D, int mode, const float* k,
                  double adt, const double* h2, int nsteps, int periodic)
{
    grid_t g = mkgrid(gx, gy, gz, D, periodic);
    long long n = grid_elems(&g);
    float c[5];
    coef_f32(D, c);
    float* v = (float*)malloc(sizeof(float) * n);
    memcpy(v, u, sizeof(float) * n);
    float* src = u;
    float* dst = v;
    for (int s = 0; s < nsteps; ++s) {
#pragma omp parallel for schedule(static)
        for (int z = 0; z < gz; ++z)
            for (int y = 0; y < gy; ++y)
                for (int x = 0; x < gx; ++x) {
                    float mx[5], px[5], my[5], py[5], mz[5], pz[5];
                    float uc = src[gidx(&g, x, y, z)];
                    for (int j = 1; j <= D; ++j) {
                        mx[j] = src[gidx(&g, x - j, y, z)]; px[j] = src[gidx(&g, x + j, y, z)];
                        my[j] = src[gidx(&g, x, y - j, z)]; py[j] = src[gidx(&g, x, y + j, z)];
                        mz[j] = src[gidx(&g, x, y, z - j)]; pz[j] = src[gidx(&g, x, y, z + j)];
                    }
                    float r;
                    if (mode == 0) {
                        float t = k[0] * lap_fma(c, D, uc, mx, px);
                        t = fmaf(k[1], lap_fma(c, D, uc, my, py), t);
                        t = fmaf(k[2], lap_fma(c, D, uc, mz, pz), t);
                        r = uc + t;
                    } else {
                        float lx = lap_ref(c, D, uc, mx, px);
                        float ly = lap_ref(c, D, uc, my, py);
                        float lz = lap_ref(c, D, uc, mz, pz);
                        double t = (double)lx / h2[0];
                        t = t + (double)ly / h2[1];
                        t = t + (double)lz / h2[2];
                        r = (float)((double)uc + adt * t);
                    }
                    dst[gidx(&g, x, y, z)] = r;
                }
        float* t = src; src = dst; dst = t;
    }
    if (src != u) memcpy(u, src, sizeof(float) * n);
    free(v);
}

/* Restatement of cpu_fdm_heat_diffuse1 (heat-3D.cpp:48-62) on one padded block:
 * region [b2r_i+1, N-2-b2r_i] per axis, z innermost and reversed, the reference
 * expression with Alpha*Dt and Dx*Dx.. as doubles. */
void orc_heat_ref1_block_step(const float* h1, float* h2, int Ngx, int Ngy, int Ngz,
                              int b2r_i, double adt, double dx2, double dy2, double dz2)
{
    long long sxy = (long long)Ngx * Ngy;
#pragma omp parallel for schedule(static)
    for (int ix = b2r_i + 1; ix <= Ngx - 2 - b2r_i; ix++)
        for (int iy = b2r_i + 1; iy <= Ngy - 2 - b2r_i; iy++)
            for (int iz = Ngz - 2 - b2r_i; iz >= b2r_i + 1; iz--) {
                long long c = iz * sxy + (long long)iy * Ngx + ix;
                float u = h1[c];
                float lx = (h1[c - 1] - 2 * u) + h1[c + 1];
                float ly = (h1[c - Ngx] - 2 * u) + h1[c + Ngx];
                float lz = (h1[c - sxy] - 2 * u) + h1[c + sxy];
                h2[c] = (float)((double)u + adt * (((double)lx / dx2 + (double)ly / dy2) +
                                                   (double)lz / dz2));
            }
}

/* heat_3D_cpu_main's R-step ping-pong (heat-3D.cpp:136-169) without its
 * copy-back (which uses the config's B2R_D for the step's D): runs R steps in
 * place on h_in using scratch. */
void orc_heat_ref1_block_run(float* h_in, float* scratch, int R, int Ngx, int Ngy, int Ngz,
                             double adt, double dx2, double dy2, double dz2)
{
    long long n = (long long)Ngx * Ngy * Ngz;
    memcpy(scratch, h_in, sizeof(float) * n);
    float* pin = h_in;
    float* pout = scratch;
    for (int i = 0; i < R; i++) {
        orc_heat_ref1_block_step(pin, pout, Ngx, Ngy, Ngz, i, adt, dx2, dy2, dz2);
        float* t = pin; pin = pout; pout = t;
    }
    if (pin != h_in) memcpy(h_in, pin, sizeof(float) * n);
}

/* ------------------------------------------------------------------------- */
/* FP64 heat adjoint (no reference counterpart)                              */
/* ------------------------------------------------------------------------- */
static inline int owned(const grid_t* g, int x, int y, int z)
{
    return x >= 0 && x < g->gx && y >= 0 && y < g->gy && z >= 0 && z < g->gz;
}

/* J = 1/2 sum_owned (u_N - tgt)^2 with u_N = nsteps FP64 heat steps from u0
 * (grids as orc_heat_f64).  g_u0 (same shape, frame left zero) = dJ/du0,
 * g_k[a] = dJ/dk_a.  All states are kept (small grids only).  The adjoint step
 * is written as the transpose of the forward gather, i.e. a scatter: every
 * owned cell q sends k_a c_j lam[q] to its owned neighbours q -/+ j e_a and
 * (1 + c_0 sum_a k_a) lam[q] to itself; frame cells are constants and
 * receive nothing. */
void orc_heat_grad_f64(const double* u0, const double* tgt, int gx, int gy, int gz, int D,
                       const double* k, int nsteps, int periodic, double* misfit,
                       double* g_u0, double* g_k)
{
    grid_t g = mkgrid(gx, gy, gz, D, periodic);
    long long n = grid_elems(&g);
    double c[5];
    coef_f64(D, c);
    double* st = (double*)malloc(sizeof(double) * n * (nsteps + 1));
    double* lam = (double*)calloc(n, sizeof(double));
    double* nl = (double*)calloc(n, sizeof(double));
    memcpy(st, u0, sizeof(double) * n);
    for (int s = 0; s < nsteps; ++s) {
        memcpy(st + (s + 1) * n, st + s * n, sizeof(double) * n);
        orc_heat_f64(st + (s + 1) * n, gx, gy, gz, D, k, 1, periodic);
    }
    double J = 0.0;
    const double* uN = st + (long long)nsteps * n;
    for (int z = 0; z < gz; ++z)
        for (int y = 0; y < gy; ++y)
            for (int x = 0; x < gx; ++x) {
                long long i = gidx(&g, x, y, z);
                double r = uN[i] - tgt[i];
                lam[i] = r;
                J += 0.5 * r * r;
            }
    double gk[3] = {0.0, 0.0, 0.0};
    const double self = 1.0 + c[0] * (k[0] + k[1] + k[2]);
    for (int s = nsteps - 1; s >= 0; --s) {
        const double* u = st + (long long)s * n;
        memset(nl, 0, sizeof(double) * n);
        for (int z = 0; z < gz; ++z)
            for (int y = 0; y < gy; ++y)
                for (int x = 0; x < gx; ++x) {
                    long long q = gidx(&g, x, y, z);
                    double uc = u[q];
                    double l[3] = {c[0] * uc, c[0] * uc, c[0] * uc};
                    for (int j = 1; j <= D; ++j) {
                        l[0] += c[j] * (u[gidx(&g, x - j, y, z)] + u[gidx(&g, x + j, y, z)]);
                        l[1] += c[j] * (u[gidx(&g, x, y - j, z)] + u[gidx(&g, x, y + j, z)]);
                        l[2] += c[j] * (u[gidx(&g, x, y, z - j)] + u[gidx(&g, x, y, z + j)]);
                    }
                    for (int a = 0; a < 3; ++a) gk[a] += lam[q] * l[a];
                    nl[q] += self * lam[q];
                    for (int j = 1; j <= D; ++j)
                        for (int sgn = -1; sgn <= 1; sgn += 2) {
                            int d = sgn * j;
                            if (periodic || owned(&g, x + d, y, z))
                                nl[gidx(&g, x + d, y, z)] += k[0] * c[j] * lam[q];
                            if (periodic || owned(&g, x, y + d, z))
                                nl[gidx(&g, x, y + d, z)] += k[1] * c[j] * lam[q];
                            if (periodic || owned(&g, x, y, z + d))
                                nl[gidx(&g, x, y, z + d)] += k[2] * c[j] * lam[q];
                        }
                }
        double* t = lam; lam = nl; nl = t;
    }
    memset(g_u0, 0, sizeof(double) * n);
    for (int z = 0; z < gz; ++z)
        for (int y = 0; y < gy; ++y)
            for (int x = 0; x < gx; ++x) {
                long long i = gidx(&g, x, y, z);
                g_u0[i] = lam[i];
            }
    for (int a = 0; a < 3; ++a) g_k[a] = gk[a];
    *misfit = J;
    free(st);
    free(lam);
    free(nl);
}

/* ------------------------------------------------------------------------- */
/* Game of Life (26-neighbour, survive 2..7, birth 5)                        */
/* ------------------------------------------------------------------------- */
static inline uint8_t gol_rule(uint8_t c, int n)
{
    if (c && n <= 1) return 0;          /* gol-3D.cu:93-94 */
    if (!c && n == 5) return 1;         /* gol-3D.cu:95-96 */
    if (c && n >= 8) return 0;          /* gol-3D.cu:97-98 */
    return c;                           /* unchanged: total rule */
}

/* Frozen: grid (gx+2)(gy+2)(gz+2) with a frame of width 1; periodic: gx*gy*gz.
 * Neighbour coordinates come from per-axis tables (wrapped for periodic). */
void orc_gol(uint8_t* u, int gx, int gy, int gz, int nsteps, int periodic)
{
    grid_t g = mkgrid(gx, gy, gz, 1, periodic);
    long long n = grid_elems(&g);
    int* tx = (int*)malloc(sizeof(int) * 3 * (gx + 2));
    int* ty = (int*)malloc(sizeof(int) * 3 * (gy + 2));
    int* tz = (int*)malloc(sizeof(int) * 3 * (gz + 2));
    /* t[3*i + (d+1)] = storage coordinate of owned coordinate i+d */
    for (int i = 0; i < gx; ++i)
        for (int d = -1; d <= 1; ++d)
            tx[3 * i + d + 1] = wrap_or_frame(i + d, gx, 1, g.periodic & 1);
    for (int i = 0; i < gy; ++i)
        for (int d = -1; d <= 1; ++d)
            ty[3 * i + d + 1] = wrap_or_frame(i + d, gy, 1, (g.periodic >> 1) & 1);
    for (int i = 0; i < gz; ++i)
        for (int d = -1; d <= 1; ++d)
            tz[3 * i + d + 1] = wrap_or_frame(i + d, gz, 1, (g.periodic >> 2) & 1);
    uint8_t* v = (uint8_t*)malloc(n);
    memcpy(v, u, n);
    uint8_t* src = u;
    uint8_t* dst = v;
    for (int s = 0; s < nsteps; ++s) {
#pragma omp parallel for schedule(static)
        for (int z = 0; z < gz; ++z)
            for (int y = 0; y < gy; ++y)
                for (int x = 0; x < gx; ++x) {
                    int cnt = 0;
                    for (int dz = 0; dz < 3; ++dz)
                        for (int dy = 0; dy < 3; ++dy) {
                            const uint8_t* row = src + ((long long)tz[3 * z + dz] * g.ny +
                                                        ty[3 * y + dy]) * g.nx;
                            cnt += row[tx[3 * x]] + row[tx[3 * x + 1]] + row[tx[3 * x + 2]];
                        }
                    long long c = ((long long)tz[3 * z + 1] * g.ny + ty[3 * y + 1]) * g.nx +
                                  tx[3 * x + 1];
                    cnt -= src[c];
                    dst[c] = gol_rule(src[c], cnt);
                }
        uint8_t* t = src; src = dst; dst = t;
    }
    if (src != u) memcpy(u, src, n);
    free(v);
    free(tx);
    free(ty);
    free(tz);
}

/* Population count of a u8 grid (a checksum used at full sizes). */
long long orc_popcount_u8(const uint8_t* u, long long n)
{
    long long s = 0;
    for (long long i = 0; i < n; ++i) s += u[i];
    return s;
}

/* ------------------------------------------------------------------------- */
/* B2R halo exchange on the library's field layout                            */
/* ------------------------------------------------------------------------- */
typedef struct {
    int nb[3], B[3], S[3], h, xoff, pitch;
    long long slice, bstride;
    int esize, periodic;
} olay_t;

static inline char* cell(const olay_t* L, char* f, int b, int x, int y, int z)
{
    long long e = (long long)b * L->bstride + L->xoff + (long long)z * L->slice +
                  (long long)y * L->pitch + x;
    return f + e * L->esize;
}

/* Neighbour block of b along dim on side (0 minus, 1 plus); -1 if none.
 * periodic: bit a set = the global faces of axis a wrap (a 2D grid, one z
 * plane, wraps x and y only; b2r_config.h:12-13). */
static int nbr(const olay_t* L, int b, int dim, int side)
{
    int c[3] = {b % L->nb[0], (b / L->nb[0]) % L->nb[1], b / (L->nb[0] * L->nb[1])};
    c[dim] += side ? 1 : -1;
    if (c[dim] < 0 || c[dim] >= L->nb[dim]) {
        if (!((L->periodic >> dim) & 1)) return -1;
        c[dim] = (c[dim] + L->nb[dim]) % L->nb[dim];
    }
    return (c[2] * L->nb[1] + c[1]) * L->nb[0] + c[0];
}

static long long msg_elems(const olay_t* L, int dim)
{
    if (dim == 0) return (long long)L->h * L->B[1] * L->B[2];
    if (dim == 1) return (long long)L->S[0] * L->h * L->B[2];
    return (long long)L->S[0] * L->S[1] * L->h;
}

/* Message element e of face slab j (0..h-1) and in-face coordinates:
 *  X: e = j*(B_Y*B_Z) + (y-h)*B_Z + (z-h)      (sub_matrix_pack_x, :209-215)
 *  Y: e = j*(S_X*B_Z) + (z-h)*S_X + x          (sub_matrix_pack_y, :242-248)
 *  Z: e = j*(S_X*S_Y) + y*S_X + x              (sub_matrix_pack_z, :274-281)
 * The minus message carries owned slabs [h, 2h), the plus message [B, B+h);
 * the receiver writes minus-side halo [0,h) from its minus neighbour's plus
 * message and plus-side halo [B+h, S) from its plus neighbour's minus message. */
static void face_loop(const olay_t* L, int dim, int pack, char* f, char* m, int b, int side)
{
    int h = L->h;
    int n0, n1;
    if (dim == 0) { n0 = L->B[1]; n1 = L->B[2]; }
    else if (dim == 1) { n0 = L->S[0]; n1 = L->B[2]; }
    else { n0 = L->S[0]; n1 = L->S[1]; }
    for (int j = 0; j < h; ++j)
        for (int q1 = 0; q1 < n1; ++q1)
            for (int q0 = 0; q0 < n0; ++q0) {
                int x, y, z;
                long long e;
                /* slab coordinate along dim: pack reads owned, unpack writes halo */
                int a;
                if (pack) a = side ? L->B[dim] + j : h + j;
                else a = side ? L->B[dim] + h + j : j;
                if (dim == 0) { x = a; y = h + q0; z = h + q1; e = (long long)j * n0 * n1 + (long long)q0 * n1 + q1; }
                else if (dim == 1) { x = q0; y = a; z = h + q1; e = (long long)j * n0 * n1 + (long long)q1 * n0 + q0; }
                else { x = q0; y = q1; z = a; e = (long long)j * n0 * n1 + (long long)q1 * n0 + q0; }
                char* c = cell(L, f, b, x, y, z);
                if (pack) memcpy(m + e * L->esize, c, L->esize);
                else memcpy(c, m + e * L->esize, L->esize);
            }
}

long long orc_msg_elems(const int* nb, const int* B, const int* S, int h, int dim)
{
    olay_t L;
    memset(&L, 0, sizeof(L));
    for (int i = 0; i < 3; ++i) { L.nb[i] = nb[i]; L.B[i] = B[i]; L.S[i] = S[i]; }
    L.h = h;
    return msg_elems(&L, dim);
}

static olay_t mklay(const int* nb, const int* B, const int* S, int h, int xoff, int pitch,
                    long long slice, long long bstride, int esize, int periodic)
{
    olay_t L;
    for (int i = 0; i < 3; ++i) { L.nb[i] = nb[i]; L.B[i] = B[i]; L.S[i] = S[i]; }
    L.h = h; L.xoff = xoff; L.pitch = pitch; L.slice = slice; L.bstride = bstride;
    L.esize = esize; L.periodic = periodic;
    return L;
}

/* Pack every block's two faces along dim (a block without a neighbour on a side
 * still packs that side; the message is simply never delivered). */
void orc_pack(const void* field, void* msgs, int dim, const int* nb, const int* B,
              const int* S, int h, int xoff, int pitch, long long slice,
              long long bstride, int esize, int periodic)
{
    olay_t L = mklay(nb, B, S, h, xoff, pitch, slice, bstride, esize, periodic);
    long long me = msg_elems(&L, dim);
    int nblk = nb[0] * nb[1] * nb[2];
    for (int b = 0; b < nblk; ++b)
        for (int side = 0; side < 2; ++side)
            face_loop(&L, dim, 1, (char*)field,
                      (char*)msgs + ((long long)b * 2 + side) * me * esize, b, side);
}

/* Unpack into every block the messages its neighbours addressed to it. */
void orc_unpack(void* field, const void* msgs, int dim, const int* nb, const int* B,
                const int* S, int h, int xoff, int pitch, long long slice,
                long long bstride, int esize, int periodic)
{
    olay_t L = mklay(nb, B, S, h, xoff, pitch, slice, bstride, esize, periodic);
    long long me = msg_elems(&L, dim);
    int nblk = nb[0] * nb[1] * nb[2];
    for (int b = 0; b < nblk; ++b)
        for (int side = 0; side < 2; ++side) {
            int src = nbr(&L, b, dim, side);
            if (src < 0) continue;
            /* my minus-side halo comes from my minus neighbour's plus message */
            char* m = (char*)msgs + ((long long)src * 2 + (1 - side)) * me * esize;
            face_loop(&L, dim, 0, (char*)field, m, b, side);
        }
}

/* boundary_updating: pack/unpack X, then Y, then Z. */
void orc_exchange(void* field, const int* nb, const int* B, const int* S, int h, int xoff,
                  int pitch, long long slice, long long bstride, int esize, int periodic)
{
    int nblk = nb[0] * nb[1] * nb[2];
    for (int dim = 0; dim < 3; ++dim) {
        olay_t L = mklay(nb, B, S, h, xoff, pitch, slice, bstride, esize, periodic);
        long long me = msg_elems(&L, dim);
        void* msgs = malloc((size_t)(me * esize * 2 * nblk));
        orc_pack(field, msgs, dim, nb, B, S, h, xoff, pitch, slice, bstride, esize, periodic);
        orc_unpack(field, msgs, dim, nb, B, S, h, xoff, pitch, slice, bstride, esize, periodic);
        free(msgs);
    }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
