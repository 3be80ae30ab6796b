// b2r_elastic.hip -- collocated 4th-order elastic velocity-stress propagation
// (the reference's eqwp model) for gfx950: forward, sources, receivers, FP64
// misfit, hand-written discrete adjoint with checkpointing.
//
// Reference: eqwp_fd4_stress eqwp.cu:341-477, eqwp_fd4_vx/vy/vz :60-331,
// eqwp_gpu_main :661-714 (stress, then vx, vy, vz per step).  First
// derivative (-1/12 (f[+2]-f[-2]) + 2/3 (f[+1]-f[-1]))/h (eqwp.cu:22-23),
// stress law of eqwp.cu:460-466 with per-cell lambda, mu and b = 1/rho.
//
// B2R geometry: one step is two dependent radius-2 stages, so a block carries
// a halo of 4R (radius 4 in the context geometry) and stage s of an R-step
// window updates [2(s+1), S-1-2(s+1)] (the reference sizes the halo R*D with
// D=2, b2r.cpp:61-64, which is too shallow for two radius-2 stages per step).
// Updates are in place: the stress stage reads v and writes sigma, the
// velocity stage reads sigma and writes v.  A thread owns one (x, y) column of
// a 64 x 4 tile and marches z; z neighbours come from register queues, x and
// y neighbours from the current plane staged in LDS with a 2-cell halo.
//
// Two forward paths run the same kernels (identical results):
//   plain: per shot and step, stress + vx + vy + vz launches and a device
//          synchronisation (the reference's structure);
//   batched: one stress and one three-component velocity launch per step for
//          all shots of a batch (a device table maps workgroups to shots of
//          different sizes), the time loop captured in HIP graphs.
// Gradients (no reference counterpart): the adjoint propagates as the forward
// scheme with -dt in scaled variables, plus kernels for the gradient terms;
// the forward states come either from per-segment recomputation (plain) or
// from running each step backwards (batched reverse sweep).
// Arithmetic is FP32 without contraction and matches the CPU oracle bit for
// bit; misfits are reduced in FP64.
#include "b2r_common.h"

#include <cstdio>
#include <cstring>
#include <vector>

enum { F_VX, F_VY, F_VZ, F_SXX, F_SYY, F_SZZ, F_SXY, F_SXZ, F_SYZ, F_LAM, F_MU, F_B, F_NUM };

// Stage-kernel tile: 64 x 8 cells, one per thread.  Against 64 x 4 the y halo
// rows are half as many per tile row (HBM bytes 1.24x -> fewer of the
// algorithmic bytes re-read): 4 % faster stress + velocity stages at 256^3.
constexpr int EL_TX = 64, EL_TY = 8;
static_assert(EL_TY >= 4 && EL_TY % 4 == 0, "the 4 halo rows come from the first 4 waves");
constexpr int EL_WG = EL_TX * EL_TY;  // threads of a tile workgroup
constexpr int EL_MAX_BATCH = 64;

struct SrcRec {
    const int* src;     // x,y,z,comp
    const float* wav;   // nt x nsrc
    int nsrc, src_nt;
    const int* rec;     // x,y,z
    float* recv;        // nt x nrec
    int nrec, rec_comp, rec_nt;
};

struct ElDev {
    KGeo g;
    float* f[F_NUM];
    float* nx[9];     // the other state set, written by the fused step kernel
    float* disp;      // optional displacement accumulator (eqwp.cu:231, 234)
    float* adj[9];    // adjoint state
    float* fw0[9];    // forward state before the step being reversed
    float* fw1[9];    // forward state after it
    float* grad[3];   // d misfit / d lambda, mu, b
    float dt;
    float ih[3];
    int ntx, nty, nzc, zc;
    int o[3], w[3];   // tiled region per axis: [o, o + w)
    int tile0;        // first workgroup of this shot in a batched launch
    int ntiles;       // tiles of this shot
    int ngrid;        // its workgroups: ntiles rounded up to a multiple of 8
    // reverse sweep of the batched gradient
    SrcRec sr;        // sources and receivers
    const float* res; // receiver residuals, rec_nt x nrec
    float* cpt;       // checkpoint slots: slot j-1 holds the state at step j*ckpt
    long long fstride;// elements per field allocation
};

struct b2r_el_shot {
    b2r_ctx* ctx;
    ElDev d;
    int t;                      // steps taken since the last reset
    int nsrc = 0, src_nt = 0;
    int* d_src = nullptr;       // x,y,z,comp per source (global coordinates)
    float* d_wav = nullptr;     // src_nt x nsrc
    int nrec = 0, rec_comp = 1, rec_nt = 0;
    int* d_rec = nullptr;       // x,y,z per receiver
    float* d_recv = nullptr;    // rec_nt x nrec samples
    void* fields[F_NUM];
    void* pp[9] = {};           // second state set of the fused step (allocated on first use)
};

constexpr float EL_C2 = -1.f / 12.f;
constexpr float EL_C1 = 2.f / 3.f;

// Field pointers come from the shot table in memory; the stencil kernels cast
// them to global memory (G) and address a block's cells by 32-bit byte offsets
// from the block's base (ldo, sto).  b2r_el_create keeps a block's field below
// 4 GiB.

// first derivative at byte offset c along the axis of byte stride s
__device__ __forceinline__ float fd4(const gfloat* __restrict__ p, uint32_t c, uint32_t s, float ih)
{
    return (EL_C2 * (ldo(p, c + 2 * s) - ldo(p, c - 2 * s)) + EL_C1 * (ldo(p, c + s) - ldo(p, c - s))) *
           ih;
}

// workgroup -> (shot, block, tile) for batched launches.  A device table of
// n > 1 shots is followed by an int array holding each shot's first tile.
// Tiles are TX x TY cells; a thread's cell lies OFF cells before the tile's
// origin plus its thread index (OFF > 0: the workgroup also covers a ring).
struct Tile {
    int shot, b, x, y, zs, ze;
};

template <int TX = EL_TX, int TY = EL_TY, int OFF = 0>
__device__ __forceinline__ Tile decode_tile(const ElDev* __restrict__ tab, int nshot)
{
    const int gt = blockIdx.x;
    int s = 0;
    if (nshot > 1) {
        // binary search of the compact first-tile array that follows the table
        const int* __restrict__ t0 = reinterpret_cast<const int*>(tab + nshot);
        int hi = nshot - 1;
        while (s < hi) {
            const int mid = (s + hi + 1) >> 1;
            if (t0[mid] <= gt) s = mid;
            else hi = mid - 1;
        }
    }
    const ElDev& d = tab[s];
    // every shot starts on a multiple of 8 workgroups, so dispatch order spreads
    // it over the 8 XCDs; xcd_tile gives each XCD a contiguous part of it.
    // Tiles past ntiles decode to a block index >= nblk and are skipped.
    int lt = xcd_tile(gt - d.tile0, d.ngrid);
    Tile T;
    T.shot = s;
    const int txi = lt % d.ntx;
    lt /= d.ntx;
    const int tyi = lt % d.nty;
    lt /= d.nty;
    const int zci = lt % d.nzc;
    T.b = lt / d.nzc;
    T.x = d.o[0] + txi * TX + (int)threadIdx.x - OFF;
    T.y = d.o[1] + tyi * TY + (int)threadIdx.y - OFF;
    T.zs = d.o[2] + zci * d.zc;
    T.ze = min(T.zs + d.zc, d.o[2] + d.w[2]);  // exclusive
    return T;
}

// z-derivative from a register queue over planes z-2..z+2 (the rounding of fd4)
__device__ __forceinline__ float fdq(const float (&q)[5], float ih)
{
    return (EL_C2 * (q[4] - q[0]) + EL_C1 * (q[3] - q[1])) * ih;
}

__device__ __forceinline__ void qshift(float (&q)[5])
{
    q[0] = q[1];
    q[1] = q[2];
    q[2] = q[3];
    q[3] = q[4];
}

// A plane of fields staged in LDS for the x and y derivatives: the tile's
// cells plus the 2-cell plus-shaped halo of fd4 -- rows y0-2, y0-1, y0+TY,
// y0+TY+1 (one cell per thread) and columns x0-2, x0-1, x0+TX, x0+TX+1 (the
// first 4*TY threads).  Cells outside the block stay 0; no updated cell reads
// them.
constexpr int EL_LX = EL_TX + 4, EL_LY = EL_TY + 4;
struct Halo {
    int ly, lx;         // this thread's cell in the LDS plane
    int hry;            // LDS row of its halo-row cell (column lx)
    int hcy, hcx;       // LDS row and column of its halo-column cell
    bool in;            // its own cell lies in the block
    bool hr_any;        // it owns a halo-row cell (the first 4 waves)
    bool hc_any;        // it owns a halo-column cell
    bool hr_ok, hc_ok;  // the halo cells lie in the block
    uint32_t hro, hco;  // their byte offsets within a plane
};

__device__ __forceinline__ Halo el_halo(const KGeo& g, int x, int y, uint32_t sy)
{
    Halo h;
    const int lane = threadIdx.x, ty = threadIdx.y, tid = ty * EL_TX + lane;
    const int x0 = x - lane, y0 = y - ty;
    h.ly = ty + 2;
    h.lx = lane + 2;
    h.in = x < g.S[0] && y < g.S[1];
    const int hr = tid / EL_TX;
    h.hr_any = EL_TY == 4 || hr < 4;
    h.hry = hr < 2 ? hr : hr < 4 ? EL_TY + hr : 0;
    const int hy = y0 - 2 + h.hry;
    h.hr_ok = h.hr_any && hy >= 0 && hy < g.S[1] && x < g.S[0];
    h.hro = h.hr_ok ? (uint32_t)hy * sy + 4u * x : 0u;
    h.hc_any = tid < 4 * EL_TY;
    const int hcc = tid % 4;
    h.hcy = tid / 4 + 2;
    h.hcx = hcc < 2 ? hcc : EL_TX + hcc;
    const int hx = x0 - 2 + h.hcx, hcy = y0 + tid / 4;
    h.hc_ok = h.hc_any && hx >= 0 && hx < g.S[0] && hcy < g.S[1];
    h.hco = h.hc_ok ? (uint32_t)hcy * sy + 4u * hx : 0u;
    return h;
}

// the halo values of field f at the plane of byte offset pz
struct HaloV {
    float r, c;
};
__device__ __forceinline__ HaloV el_halo_load(const gfloat* __restrict__ f, const Halo& h,
                                              uint32_t pz)
{
    HaloV v;
    v.r = h.hr_ok ? ldo(f, pz + h.hro) : 0.0f;
    v.c = h.hc_ok ? ldo(f, pz + h.hco) : 0.0f;
    return v;
}

__device__ __forceinline__ void el_stage(float (&p)[EL_LY][EL_LX], const Halo& h, float center,
                                         HaloV hv)
{
    p[h.ly][h.lx] = center;
    if (h.hr_any) p[h.hry][h.lx] = hv.r;
    if (h.hc_any) p[h.hcy][h.hcx] = hv.c;
}

// fd4 along x and y of a staged plane at this thread's cell
__device__ __forceinline__ float fdx(const float (&p)[EL_LY][EL_LX], const Halo& h, float ih)
{
    return (EL_C2 * (p[h.ly][h.lx + 2] - p[h.ly][h.lx - 2]) +
            EL_C1 * (p[h.ly][h.lx + 1] - p[h.ly][h.lx - 1])) * ih;
}
__device__ __forceinline__ float fdy(const float (&p)[EL_LY][EL_LX], const Halo& h, float ih)
{
    return (EL_C2 * (p[h.ly + 2][h.lx] - p[h.ly - 2][h.lx]) +
            EL_C1 * (p[h.ly + 1][h.lx] - p[h.ly - 1][h.lx])) * ih;
}

// ---------------------------------------------------------------- forward
// stress stage s = 2i: sigma += dt*C:grad(v) over update_box(radius 2, s).
// v of the current plane is staged in LDS for the x and y derivatives; z
// neighbours come from register queues.
__global__ __launch_bounds__(EL_WG) void el_stress_kernel(const ElDev* __restrict__ tab, int nshot,
                                                        int stage)
{
    __shared__ float sv[3][EL_LY][EL_LX];
    const Tile T = decode_tile(tab, nshot);
    const ElDev& d = tab[T.shot];
    if (T.b >= d.g.nblk) return;  // padding tile (uniform)
    const KGeo& g = d.g;
    const Box box = update_box(g, T.b, 2, stage);
    const int z0 = max(T.zs, box.lo[2]), z1 = min(T.ze - 1, box.hi[2]);
    if (z0 > z1) return;  // uniform
    const int x = T.x, y = T.y;
    const bool upd = x >= box.lo[0] && x <= box.hi[0] && y >= box.lo[1] && y <= box.hi[1];
    const long long boff = (long long)T.b * g.bstride + g.xoff;
    const gfloat* __restrict__ vx = G(d.f[F_VX] + boff);
    const gfloat* __restrict__ vy = G(d.f[F_VY] + boff);
    const gfloat* __restrict__ vz = G(d.f[F_VZ] + boff);
    const gfloat* __restrict__ lam = G(d.f[F_LAM] + boff);
    const gfloat* __restrict__ mu = G(d.f[F_MU] + boff);
    gfloat* __restrict__ sxx = G(d.f[F_SXX] + boff);
    gfloat* __restrict__ syy = G(d.f[F_SYY] + boff);
    gfloat* __restrict__ szz = G(d.f[F_SZZ] + boff);
    gfloat* __restrict__ sxy = G(d.f[F_SXY] + boff);
    gfloat* __restrict__ sxz = G(d.f[F_SXZ] + boff);
    gfloat* __restrict__ syz = G(d.f[F_SYZ] + boff);
    const uint32_t sy = 4u * g.pitch, sz = 4u * (uint32_t)g.slice;  // byte strides
    const float dt = d.dt, i0 = d.ih[0], i1 = d.ih[1], i2 = d.ih[2];
    const Halo h = el_halo(g, x, y, sy);
    uint32_t c = (uint32_t)z0 * sz + (h.in ? (uint32_t)y * sy + 4u * x : 0u);
    float qx[5], qy[5], qz[5];  // v along z, planes z-2..z+2
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t q = c - 2 * sz + k * sz;
        qx[k] = h.in ? ldo(vx, q) : 0.0f;
        qy[k] = h.in ? ldo(vy, q) : 0.0f;
        qz[k] = h.in ? ldo(vz, q) : 0.0f;
    }
    for (int z = z0; z <= z1; ++z, c += sz) {
        qx[4] = h.in ? ldo(vx, c + 2 * sz) : 0.0f;
        qy[4] = h.in ? ldo(vy, c + 2 * sz) : 0.0f;
        qz[4] = h.in ? ldo(vz, c + 2 * sz) : 0.0f;
        const uint32_t pz = (uint32_t)z * sz;
        const HaloV hx = el_halo_load(vx, h, pz), hy = el_halo_load(vy, h, pz),
                    hz = el_halo_load(vz, h, pz);
        __syncthreads();  // the previous plane's reads are done
        el_stage(sv[0], h, qx[2], hx);
        el_stage(sv[1], h, qy[2], hy);
        el_stage(sv[2], h, qz[2], hz);
        __syncthreads();
        if (upd) {
            const float exx = fdx(sv[0], h, i0);
            const float eyy = fdy(sv[1], h, i1);
            const float ezz = fdq(qz, i2);
            const float dxy = fdy(sv[0], h, i1) + fdx(sv[1], h, i0);
            const float dxz = fdq(qx, i2) + fdx(sv[2], h, i0);
            const float dyz = fdq(qy, i2) + fdy(sv[2], h, i1);
            const float l = ldo(lam, c), m = ldo(mu, c);
            const float l2m = l + 2.0f * m;
            sto(sxx, c, ldo(sxx, c) + dt * (l2m * exx + l * (eyy + ezz)));
            sto(syy, c, ldo(syy, c) + dt * (l2m * eyy + l * (exx + ezz)));
            sto(szz, c, ldo(szz, c) + dt * (l2m * ezz + l * (exx + eyy)));
            sto(sxy, c, ldo(sxy, c) + dt * (m * dxy));
            sto(sxz, c, ldo(sxz, c) + dt * (m * dxz));
            sto(syz, c, ldo(syz, c) + dt * (m * dyz));
        }
        qshift(qx);
        qshift(qy);
        qshift(qz);
    }
}

// velocity stage s = 2i+1 for the components in MASK (1 x, 2 y, 4 z).
// The stresses with x or y derivatives are staged in LDS per plane; z
// neighbours come from register queues rounded exactly as direct loads.
template <int MASK>
__global__ __launch_bounds__(EL_WG) void el_velocity_kernel(const ElDev* __restrict__ tab, int nshot,
                                                          int stage)
{
    // staged planes: 0 sxx (vx), 1 syy (vy), 2 sxy (vx, vy), 3 sxz (vz), 4 syz (vz)
    constexpr bool ST[5] = {(MASK & 1) != 0, (MASK & 2) != 0, (MASK & 3) != 0, (MASK & 4) != 0,
                            (MASK & 4) != 0};
    __shared__ float sv[5][EL_LY][EL_LX];
    const Tile T = decode_tile(tab, nshot);
    const ElDev& d = tab[T.shot];
    if (T.b >= d.g.nblk) return;  // padding tile (uniform)
    const KGeo& g = d.g;
    const Box box = update_box(g, T.b, 2, stage);
    const int z0 = max(T.zs, box.lo[2]), z1 = min(T.ze - 1, box.hi[2]);
    if (z0 > z1) return;  // uniform
    const int x = T.x, y = T.y;
    const bool upd = x >= box.lo[0] && x <= box.hi[0] && y >= box.lo[1] && y <= box.hi[1];
    const long long boff = (long long)T.b * g.bstride + g.xoff;
    const gfloat* __restrict__ sf[5] = {G(d.f[F_SXX] + boff), G(d.f[F_SYY] + boff),
                                        G(d.f[F_SXY] + boff), G(d.f[F_SXZ] + boff),
                                        G(d.f[F_SYZ] + boff)};
    const gfloat* __restrict__ szz = G(d.f[F_SZZ] + boff);
    const gfloat* __restrict__ bb = G(d.f[F_B] + boff);
    gfloat* __restrict__ vx = G(d.f[F_VX] + boff);
    gfloat* __restrict__ vy = G(d.f[F_VY] + boff);
    gfloat* __restrict__ vz = G(d.f[F_VZ] + boff);
    gfloat* __restrict__ disp = d.disp ? G(d.disp + boff) : nullptr;
    const uint32_t sy = 4u * g.pitch, sz = 4u * (uint32_t)g.slice;  // byte strides
    const float dt = d.dt, i0 = d.ih[0], i1 = d.ih[1], i2 = d.ih[2];
    const Halo h = el_halo(g, x, y, sy);
    uint32_t c = (uint32_t)z0 * sz + (h.in ? (uint32_t)y * sy + 4u * x : 0u);
    // z-derivative fields (sxz for vx, syz for vy, szz for vz), planes z-2..z+2
    float qa[5], qb[5], qc[5];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t q = c - 2 * sz + k * sz;
        if (MASK & 1) qa[k] = h.in ? ldo(sf[3], q) : 0.0f;
        if (MASK & 2) qb[k] = h.in ? ldo(sf[4], q) : 0.0f;
        if (MASK & 4) qc[k] = h.in ? ldo(szz, q) : 0.0f;
    }
    for (int z = z0; z <= z1; ++z, c += sz) {
        const uint32_t q = c + 2 * sz;
        if (MASK & 1) qa[4] = h.in ? ldo(sf[3], q) : 0.0f;
        if (MASK & 2) qb[4] = h.in ? ldo(sf[4], q) : 0.0f;
        if (MASK & 4) qc[4] = h.in ? ldo(szz, q) : 0.0f;
        const uint32_t pz = (uint32_t)z * sz;
        float cv[5];
        HaloV hv[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            if (!ST[k]) continue;
            if (k == 3 && (MASK & 1)) cv[k] = qa[2];
            else if (k == 4 && (MASK & 2)) cv[k] = qb[2];
            else cv[k] = h.in ? ldo(sf[k], c) : 0.0f;
            hv[k] = el_halo_load(sf[k], h, pz);
        }
        __syncthreads();  // the previous plane's reads are done
#pragma unroll
        for (int k = 0; k < 5; ++k)
            if (ST[k]) el_stage(sv[k], h, cv[k], hv[k]);
        __syncthreads();
        if (upd) {
            const float tb = dt * ldo(bb, c);
            if (MASK & 1) {
                const float ax = (fdx(sv[0], h, i0) + fdy(sv[2], h, i1)) + fdq(qa, i2);
                sto(vx, c, ldo(vx, c) + tb * ax);
            }
            if (MASK & 2) {
                const float ay = (fdx(sv[2], h, i0) + fdy(sv[1], h, i1)) + fdq(qb, i2);
                sto(vy, c, ldo(vy, c) + tb * ay);
                if (disp) sto(disp, c, ldo(disp, c) + (dt * tb) * ay);
            }
            if (MASK & 4) {
                const float az = (fdx(sv[3], h, i0) + fdy(sv[4], h, i1)) + fdq(qc, i2);
                sto(vz, c, ldo(vz, c) + tb * az);
            }
        }
        if (MASK & 1) qshift(qa);
        if (MASK & 2) qshift(qb);
        if (MASK & 4) qshift(qc);
    }
}

// ---------------------------------------------------------------- fused step
// One whole step (stress stage 2i, then velocity stage 2i+1) in one launch,
// from the state in d.f into the other state set d.nx.  A workgroup of
// 64 x 16 threads, one cell each, covers a 60 x 12 tile of velocity cells and
// the 2-cell ring around it, and marches z: at plane p it forms the new
// stresses of all its cells (the ring's are the ones the velocity stencil of
// the tile needs, recomputed here instead of written and read back), at plane
// p-2 the new velocities of the tile.  Because a step reads one state set and
// writes the other, neighbouring workgroups, whose rings overlap this tile,
// never see a half-updated cell.  Per cell and step 9 state and 3 medium
// fields are read and 9 written (84 B), against 120 B for the two in-place
// stages.  The results equal el_stress_kernel followed by
// el_velocity_kernel<7> bit for bit (same expressions and rounding); cells
// outside the stress box keep their stresses as there.  New stresses are
// written on the velocity box only: the later stages of the window read them
// nowhere else, and the exchange after the window refreshes every halo cell.
constexpr int EF_TX = 64, EF_TY = 16;                // threads of a fused workgroup
constexpr int EF_IX = EF_TX - 4, EF_IY = EF_TY - 4;  // its tile of velocity cells
constexpr int EF_WG = EF_TX * EF_TY;
constexpr int EF_LX = EF_TX + 4, EF_LY = EF_TY + 4;  // staged v plane with its halo

template <int NY, int NX>
__device__ __forceinline__ float fdx_at(const float (&p)[NY][NX], int y, int x, float ih)
{
    return (EL_C2 * (p[y][x + 2] - p[y][x - 2]) + EL_C1 * (p[y][x + 1] - p[y][x - 1])) * ih;
}
template <int NY, int NX>
__device__ __forceinline__ float fdy_at(const float (&p)[NY][NX], int y, int x, float ih)
{
    return (EL_C2 * (p[y + 2][x] - p[y - 2][x]) + EL_C1 * (p[y + 1][x] - p[y - 1][x])) * ih;
}

__global__ __launch_bounds__(EF_WG) void el_step_kernel(const ElDev* __restrict__ tab, int nshot,
                                                      int i)
{
    // two buffers each, alternating between planes, so one barrier per plane
    // separates the staging of a plane from the reads of the previous one
    __shared__ float sv[2][3][EF_LY][EF_LX];  // v at plane p with the 2-cell plus halo
    __shared__ float ss[2][5][EF_TY][EF_TX];  // new sxx, syy, sxy, sxz, syz at plane p-2
    const Tile T = decode_tile<EF_IX, EF_IY, 2>(tab, nshot);
    const ElDev& d = tab[T.shot];
    if (T.b >= d.g.nblk) return;  // padding tile (uniform)
    const KGeo& g = d.g;
    const Box bs = update_box(g, T.b, 2, 2 * i), bv = update_box(g, T.b, 2, 2 * i + 1);
    const int k0 = max(T.zs, bv.lo[2]), k1 = min(T.ze - 1, bv.hi[2]);
    if (k0 > k1) return;  // uniform
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * EF_TX + tx;
    // the tiled region starts at 4 or deeper, so x, y >= 2 and every stencil
    // of an updated cell stays in the block; planes k0-4 .. k1+4 lie in it too
    const int x = T.x, y = T.y;
    const bool in = x < g.S[0] && y < g.S[1];
    const bool s_xy = x >= bs.lo[0] && x <= bs.hi[0] && y >= bs.lo[1] && y <= bs.hi[1];
    const bool v_xy = tx >= 2 && tx < EF_TX - 2 && ty >= 2 && ty < EF_TY - 2 && x >= bv.lo[0] &&
                      x <= bv.hi[0] && y >= bv.lo[1] && y <= bv.hi[1];
    const long long boff = (long long)T.b * g.bstride + g.xoff;
    const gfloat* __restrict__ fv[3] = {G(d.f[F_VX] + boff), G(d.f[F_VY] + boff),
                                        G(d.f[F_VZ] + boff)};
    const gfloat* __restrict__ fs[6] = {G(d.f[F_SXX] + boff), G(d.f[F_SYY] + boff),
                                        G(d.f[F_SZZ] + boff), G(d.f[F_SXY] + boff),
                                        G(d.f[F_SXZ] + boff), G(d.f[F_SYZ] + boff)};
    const gfloat* __restrict__ lam = G(d.f[F_LAM] + boff);
    const gfloat* __restrict__ mu = G(d.f[F_MU] + boff);
    const gfloat* __restrict__ bb = G(d.f[F_B] + boff);
    gfloat* __restrict__ ov[3] = {G(d.nx[F_VX] + boff), G(d.nx[F_VY] + boff), G(d.nx[F_VZ] + boff)};
    gfloat* __restrict__ os[6] = {G(d.nx[F_SXX] + boff), G(d.nx[F_SYY] + boff),
                                  G(d.nx[F_SZZ] + boff), G(d.nx[F_SXY] + boff),
                                  G(d.nx[F_SXZ] + boff), G(d.nx[F_SYZ] + boff)};
    gfloat* __restrict__ disp = d.disp ? G(d.disp + boff) : nullptr;
    const uint32_t sy = 4u * g.pitch, sz = 4u * (uint32_t)g.slice;  // byte strides
    const float dt = d.dt, i0 = d.ih[0], i1 = d.ih[1], i2 = d.ih[2];
    // halo of the staged v plane: waves 0-3 hold rows y0-2, y0-1, y0+16, y0+17
    // of the workgroup's cells, wave 4 columns x0-2, x0-1, x0+64, x0+65
    const bool h_any = tid < 4 * EF_TX + 4 * EF_TY;
    int hly, hlx, hx, hy;
    if (tid < 4 * EF_TX) {
        const int r = tid / EF_TX;
        hly = r < 2 ? r : EF_TY + r;
        hlx = tx + 2;
        hx = x;
        hy = y - ty - 2 + hly;
    } else {
        const int j = tid - 4 * EF_TX, cc = j & 3;
        hlx = cc < 2 ? cc : EF_TX + cc;
        hly = (j >> 2) + 2;
        hx = x - tx - 2 + hlx;
        hy = y - ty - 2 + hly;
    }
    const bool h_ok = h_any && hx >= 0 && hx < g.S[0] && hy >= 0 && hy < g.S[1];
    const uint32_t ho = h_ok ? (uint32_t)hy * sy + 4u * hx : 0u;
    const uint32_t cxy = in ? (uint32_t)y * sy + 4u * x : 0u;
    // register queues: v over planes p-2..p+2; new sxz, syz, szz over p-4..p;
    // new sxx, syy, sxy over p-2..p
    float vq[3][5], zq[3][5], xq[3][3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
#pragma unroll
        for (int j = 0; j < 4; ++j) vq[f][j] = ldo(fv[f], (uint32_t)(k0 - 4 + j) * sz + cxy);
#pragma unroll
        for (int j = 0; j < 5; ++j) zq[f][j] = 0.0f;
#pragma unroll
        for (int j = 0; j < 3; ++j) xq[f][j] = 0.0f;
    }
    // Every load is issued one plane ahead of its use (v of plane p+3, the
    // halo, old stresses and medium of plane p+1, b of plane p-1) into one of
    // two register sets that alternate between planes, so no load in flight
    // is moved between registers.  Loads are unconditional: a cell outside
    // the block reads the block's cell (0, 0) of its plane instead (cxy, ho =
    // 0), a value no updated cell uses.
    struct Pre {
        float v[3], h[3], s[6], l, m, b;
    };
    auto load = [&](Pre& n, int p) {  // the set consumed at plane p
        const uint32_t pz = (uint32_t)p * sz, c = pz + cxy;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            n.v[f] = ldo(fv[f], c + 2 * sz);
            n.h[f] = h_any ? ldo(fv[f], pz + ho) : 0.0f;  // uniform: waves 0-4
        }
#pragma unroll
        for (int f = 0; f < 6; ++f) n.s[f] = ldo(fs[f], c);
        n.l = ldo(lam, c);
        n.m = ldo(mu, c);
        n.b = ldo(bb, c - 2 * sz);
    };
    auto plane = [&](int p, const Pre& cur, Pre& nxt, int buf) {
        load(nxt, min(p + 1, k1 + 2));
        const uint32_t ck = (uint32_t)(p - 2) * sz + cxy;
#pragma unroll
        for (int f = 0; f < 3; ++f) vq[f][4] = cur.v[f];
        // stage v of plane p and the new stresses of plane k = p-2 (the reads
        // of this buffer, two planes back, ended before the last barrier)
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            sv[buf][f][ty + 2][tx + 2] = vq[f][2];
            if (h_any) sv[buf][f][hly][hlx] = cur.h[f];
        }
        ss[buf][0][ty][tx] = xq[0][0];
        ss[buf][1][ty][tx] = xq[1][0];
        ss[buf][2][ty][tx] = xq[2][0];
        ss[buf][3][ty][tx] = zq[0][2];
        ss[buf][4][ty][tx] = zq[1][2];
        __syncthreads();
        // new stresses of plane p (el_stress_kernel's expressions), computed
        // by every thread and kept where the cell is in the stress box
        const bool upd = s_xy && p >= bs.lo[2] && p <= bs.hi[2];
        float n[6];
        {
            const auto& V = sv[buf];
            const int ly = ty + 2, lx = tx + 2;
            const float l = cur.l, m = cur.m;
            const float exx = fdx_at(V[0], ly, lx, i0);
            const float eyy = fdy_at(V[1], ly, lx, i1);
            const float ezz = fdq(vq[2], i2);
            const float dxy = fdy_at(V[0], ly, lx, i1) + fdx_at(V[1], ly, lx, i0);
            const float dxz = fdq(vq[0], i2) + fdx_at(V[2], ly, lx, i0);
            const float dyz = fdq(vq[1], i2) + fdy_at(V[2], ly, lx, i1);
            const float l2m = l + 2.0f * m;
            n[0] = cur.s[0] + dt * (l2m * exx + l * (eyy + ezz));
            n[1] = cur.s[1] + dt * (l2m * eyy + l * (exx + ezz));
            n[2] = cur.s[2] + dt * (l2m * ezz + l * (exx + eyy));
            n[3] = cur.s[3] + dt * (m * dxy);
            n[4] = cur.s[4] + dt * (m * dxz);
            n[5] = cur.s[5] + dt * (m * dyz);
        }
#pragma unroll
        for (int f = 0; f < 6; ++f) n[f] = upd ? n[f] : cur.s[f];
        xq[0][2] = n[0];
        xq[1][2] = n[1];
        xq[2][2] = n[3];
        zq[0][4] = n[4];
        zq[1][4] = n[5];
        zq[2][4] = n[2];
        // new velocities of plane k (el_velocity_kernel's expressions)
        if (p - 2 >= k0 && p - 2 <= k1 && v_xy) {
            const auto& S = ss[buf];
            const float tb = dt * cur.b;
            const float ax = (fdx_at(S[0], ty, tx, i0) + fdy_at(S[2], ty, tx, i1)) + fdq(zq[0], i2);
            const float ay = (fdx_at(S[2], ty, tx, i0) + fdy_at(S[1], ty, tx, i1)) + fdq(zq[1], i2);
            const float az = (fdx_at(S[3], ty, tx, i0) + fdy_at(S[4], ty, tx, i1)) + fdq(zq[2], i2);
            sto(ov[0], ck, vq[0][0] + tb * ax);
            sto(ov[1], ck, vq[1][0] + tb * ay);
            sto(ov[2], ck, vq[2][0] + tb * az);
            if (disp) sto(disp, ck, ldo(disp, ck) + (dt * tb) * ay);
            sto(os[0], ck, xq[0][0]);
            sto(os[1], ck, xq[1][0]);
            sto(os[2], ck, zq[2][2]);
            sto(os[3], ck, xq[2][0]);
            sto(os[4], ck, zq[0][2]);
            sto(os[5], ck, zq[1][2]);
        }
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            qshift(vq[f]);
            qshift(zq[f]);
            xq[f][0] = xq[f][1];
            xq[f][1] = xq[f][2];
        }
    };
    // planes in pairs; an odd count ends with a plane k1+3 that reloads
    // plane k1+2 and stores nothing
    Pre pa, pb;
    load(pa, k0 - 2);
    for (int p = k0 - 2; p <= k1 + 2; p += 2) {
        plane(p, pa, pb, 0);
        plane(p + 1, pb, pa, 1);
    }
}

// Point forces after the velocity stage of inner step i: every block whose
// stage box holds the source cell adds dt*b*f there, so redundant halo
// computation between exchanges stays consistent with the owner.  One launch
// serves a table of shots: blockIdx.y is the shot, tcount[shot] its step.
__global__ void el_inject_kernel(const ElDev* __restrict__ tab, int stage,
                                 const int* __restrict__ tcount)
{
    const ElDev& d = tab[blockIdx.y];
    const SrcRec& sr = d.sr;
    const int t = tcount[blockIdx.y];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= sr.nsrc * d.g.nblk || t >= sr.src_nt) return;
    const int s = idx / d.g.nblk, b = idx % d.g.nblk;
    int bc[3];
    block_coords(d.g, b, bc);
    const int* p = sr.src + 4 * s;
    const Box box = update_box(d.g, b, 2, stage);
    const float w = sr.wav[(long long)t * sr.nsrc + s];
    float* v = d.f[p[3]];
    // with periodic faces the source cell also has images one period away,
    // which can lie in this block's halo (or, for a single block along an
    // axis, in its box twice); the redundant halo stages see every image
    const int px = d.g.periodic & 1, py = (d.g.periodic >> 1) & 1, pz = (d.g.periodic >> 2) & 1;
    for (int kz = -pz; kz <= pz; ++kz)
        for (int ky = -py; ky <= py; ++ky)
            for (int kx = -px; kx <= px; ++kx) {
                const int sh[3] = {kx, ky, kz};
                int lc[3];
                bool in = true;
                for (int a = 0; a < 3; ++a) {
                    lc[a] = p[a] + sh[a] * d.g.nb[a] * d.g.B[a] - bc[a] * d.g.B[a] + d.g.h;
                    in = in && lc[a] >= box.lo[a] && lc[a] <= box.hi[a];
                }
                if (!in) continue;
                const long long c = cell_index(d.g, b, lc[0], lc[1], lc[2]);
                const float tb = d.dt * d.f[F_B][c];
                v[c] = v[c] + tb * w;
                // a y force also enters the displacement accumulator, as
                // Dt*Dt*inv_rho*(div sigma + f.y) does (eqwp.cu:229-235)
                if (d.disp && p[3] == F_VY) d.disp[c] = d.disp[c] + (d.dt * tb) * w;
            }
}

// Receivers sample their component from the owning block after the step.
__global__ void el_record_kernel(const ElDev* __restrict__ tab, const int* __restrict__ tcount)
{
    const ElDev& d = tab[blockIdx.y];
    const SrcRec& sr = d.sr;
    const int t = tcount[blockIdx.y];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= sr.nrec || t >= sr.rec_nt) return;
    const int* p = sr.rec + 3 * r;
    int bc[3], lc[3];
    for (int a = 0; a < 3; ++a) {
        bc[a] = p[a] / d.g.B[a];
        lc[a] = p[a] - bc[a] * d.g.B[a] + d.g.h;
    }
    const int b = (bc[2] * d.g.nb[1] + bc[1]) * d.g.nb[0] + bc[0];
    sr.recv[(long long)t * sr.nrec + r] = d.f[sr.rec_comp][cell_index(d.g, b, lc[0], lc[1], lc[2])];
}

// step counters of n shots += delta
__global__ void el_tick_kernel(int* tcount, int n, int delta)
{
    const int i = threadIdx.x;
    if (i < n) tcount[i] += delta;
}

// ---------------------------------------------------------------- adjoint
// Adjoint of one forward step at R = 1, every global face frozen or periodic.
// With E the strain-rate operator and M the force divergence (M = -E^T on the
// zero-framed or the periodic grid, D^T = -D), a step is
//   sigma' = sigma + dt C E v,   v' = v + dt B M sigma'      (B = diag(b))
// and its transpose, in the variables w = B vh and p = C sh (vh, sh the
// adjoints of v and sigma), is the same scheme run with -dt:
//   p <- p - dt C E w,           w <- w - dt B M p,
// so the adjoint propagation reuses el_stress_kernel and el_velocity_kernel
// on a table whose fields are (w, p) and whose dt is negated.  The kernels
// below add the gradients,
//   g_b      += dt (w/b) . M sigma'          (sigma' = stresses after the stress stage)
//   g_lambda += dt tr(sh) tr(e),  g_mu += dt (2 sh_ii e_ii + sh_ij d_ij),
//   sh = C^-1 p,  e = E v                     (v = velocity before the step)
// and serve two algorithms:
//   REV = false (plain): sigma' and v come from stored copies (fw1, fw0);
//   REV = true (reverse sweep): d.f holds the state after the step and is
//     reconstructed backwards in place, the scheme being time-reversible:
//     v = v' - dt b M sigma' - dt b f(t) at the sources, then
//     sigma = sigma' - dt C E v.  Nothing is stored between checkpoints; the
//     checkpoint reset every ckpt steps bounds the rounding drift.
// The kernels update the boxes of the forward's stages: the velocity stage's
// gradient and reconstruction run on the stress stage's box (the owned cells
// and, toward neighbouring blocks, a 2-cell ring, whose rebuilt velocities the
// stress stage's stencil reads), the stress stage's on the velocity stage's
// (the owned cells); gradients are accumulated on owned cells only.  With
// several blocks or periodic faces the reverse sweep exchanges the state and
// the adjoint before every step, as the forward exchanges before it.
// Both kernels write cell-local values only and read with a stencil only
// fields they do not write, so they run in place; z neighbours come from
// register queues rounded exactly as direct loads.

// velocity stage: a = M sigma' (rounded as the forward velocity stage);
// g_b += dt (w/b).a;  REV: v -= dt b a.  The stresses are staged in LDS for
// the x and y derivatives as in el_velocity_kernel.
template <bool REV>
__global__ __launch_bounds__(EL_WG) void el_adj_vel_kernel(const ElDev* __restrict__ tab, int nshot)
{
    __shared__ float sv[5][EL_LY][EL_LX];  // sxx, syy, sxy, sxz, syz
    const Tile T = decode_tile(tab, nshot);
    const ElDev& d = tab[T.shot];
    if (T.b >= d.g.nblk) return;  // padding tile (uniform)
    const KGeo& g = d.g;
    const int hh = g.h, x = T.x, y = T.y;
    // the box of the forward's stress stage: the owned cells and a 2-cell ring
    // toward neighbouring blocks (where REV rebuilds the velocities that the
    // stress stage's stencil reads); the gradient takes owned cells only
    const Box box = update_box(g, T.b, 2, 0);
    const int z0 = max(T.zs, box.lo[2]), z1 = min(T.ze - 1, box.hi[2]);
    if (z0 > z1) return;  // uniform
    const bool upd = x >= box.lo[0] && x <= box.hi[0] && y >= box.lo[1] && y <= box.hi[1];
    const bool own_xy = x >= hh && x < hh + g.B[0] && y >= hh && y < hh + g.B[1];
    const long long boff = (long long)T.b * g.bstride + g.xoff;
    const uint32_t sy = 4u * g.pitch, sz = 4u * (uint32_t)g.slice;  // byte strides
    const float dt = d.dt, i0 = d.ih[0], i1 = d.ih[1], i2 = d.ih[2];
    float* const* sg = REV ? d.f + F_SXX : d.fw1 + F_SXX;
    const gfloat* __restrict__ sf[5] = {G(sg[0] + boff), G(sg[1] + boff), G(sg[3] + boff),
                                        G(sg[4] + boff), G(sg[5] + boff)};
    const gfloat* __restrict__ szz = G(sg[2] + boff);
    const gfloat* __restrict__ b = G(d.f[F_B] + boff);
    const gfloat* __restrict__ wx = G(d.adj[0] + boff);
    const gfloat* __restrict__ wy = G(d.adj[1] + boff);
    const gfloat* __restrict__ wz = G(d.adj[2] + boff);
    gfloat* __restrict__ gb = G(d.grad[2] + boff);
    gfloat* __restrict__ vx = G(d.f[F_VX] + boff);
    gfloat* __restrict__ vy = G(d.f[F_VY] + boff);
    gfloat* __restrict__ vz = G(d.f[F_VZ] + boff);
    const Halo h = el_halo(g, x, y, sy);
    uint32_t c = (uint32_t)z0 * sz + (h.in ? (uint32_t)y * sy + 4u * x : 0u);
    float qa[5], qb[5], qc[5];  // sxz, syz, szz along z
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t q = c - 2 * sz + k * sz;
        qa[k] = h.in ? ldo(sf[3], q) : 0.0f;
        qb[k] = h.in ? ldo(sf[4], q) : 0.0f;
        qc[k] = h.in ? ldo(szz, q) : 0.0f;
    }
    for (int z = z0; z <= z1; ++z, c += sz) {
        const uint32_t q = c + 2 * sz;
        qa[4] = h.in ? ldo(sf[3], q) : 0.0f;
        qb[4] = h.in ? ldo(sf[4], q) : 0.0f;
        qc[4] = h.in ? ldo(szz, q) : 0.0f;
        const uint32_t pz = (uint32_t)z * sz;
        float cv[5];
        HaloV hv[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            cv[k] = k == 3 ? qa[2] : k == 4 ? qb[2] : (h.in ? ldo(sf[k], c) : 0.0f);
            hv[k] = el_halo_load(sf[k], h, pz);
        }
        __syncthreads();  // the previous plane's reads are done
#pragma unroll
        for (int k = 0; k < 5; ++k) el_stage(sv[k], h, cv[k], hv[k]);
        __syncthreads();
        if (upd) {
            const float ax = (fdx(sv[0], h, i0) + fdy(sv[2], h, i1)) + fdq(qa, i2);
            const float ay = (fdx(sv[2], h, i0) + fdy(sv[1], h, i1)) + fdq(qb, i2);
            const float az = (fdx(sv[3], h, i0) + fdy(sv[4], h, i1)) + fdq(qc, i2);
            const float bc = ldo(b, c);
            if (own_xy && z >= hh && z < hh + g.B[2])
                sto(gb, c,
                    ldo(gb, c) + dt * ((ldo(wx, c) * ax + ldo(wy, c) * ay + ldo(wz, c) * az) / bc));
            if (REV) {
                const float tb = dt * bc;
                sto(vx, c, ldo(vx, c) - tb * ax);
                sto(vy, c, ldo(vy, c) - tb * ay);
                sto(vz, c, ldo(vz, c) - tb * az);
            }
        }
        qshift(qa);
        qshift(qb);
        qshift(qc);
    }
}

// stress stage: e = E v (rounded as the forward stress stage), sh = C^-1 p;
// g_lambda, g_mu += ...;  REV: sigma -= dt C e.  v is staged in LDS for the x
// and y derivatives as in el_stress_kernel.
template <bool REV>
__global__ __launch_bounds__(EL_WG) void el_adj_str_kernel(const ElDev* __restrict__ tab, int nshot)
{
    __shared__ float sv[3][EL_LY][EL_LX];
    const Tile T = decode_tile(tab, nshot);
    const ElDev& d = tab[T.shot];
    if (T.b >= d.g.nblk) return;  // padding tile (uniform)
    const KGeo& g = d.g;
    const int x = T.x, y = T.y;
    // the box of the forward's velocity stage: at R = 1 the owned cells
    const Box box = update_box(g, T.b, 2, 1);
    const int z0 = max(T.zs, box.lo[2]), z1 = min(T.ze - 1, box.hi[2]);
    if (z0 > z1) return;  // uniform
    const bool upd = x >= box.lo[0] && x <= box.hi[0] && y >= box.lo[1] && y <= box.hi[1];
    const long long boff = (long long)T.b * g.bstride + g.xoff;
    const uint32_t sy = 4u * g.pitch, sz = 4u * (uint32_t)g.slice;  // byte strides
    const float dt = d.dt, i0 = d.ih[0], i1 = d.ih[1], i2 = d.ih[2];
    float* const* vv = REV ? d.f : d.fw0;
    const gfloat* __restrict__ vx = G(vv[0] + boff);
    const gfloat* __restrict__ vy = G(vv[1] + boff);
    const gfloat* __restrict__ vz = G(vv[2] + boff);
    const gfloat* __restrict__ lam = G(d.f[F_LAM] + boff);
    const gfloat* __restrict__ mu = G(d.f[F_MU] + boff);
    const Halo h = el_halo(g, x, y, sy);
    uint32_t c = (uint32_t)z0 * sz + (h.in ? (uint32_t)y * sy + 4u * x : 0u);
    float qx[5], qy[5], qz[5];  // v along z
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t q = c - 2 * sz + k * sz;
        qx[k] = h.in ? ldo(vx, q) : 0.0f;
        qy[k] = h.in ? ldo(vy, q) : 0.0f;
        qz[k] = h.in ? ldo(vz, q) : 0.0f;
    }
    for (int z = z0; z <= z1; ++z, c += sz) {
        const uint32_t q = c + 2 * sz;
        qx[4] = h.in ? ldo(vx, q) : 0.0f;
        qy[4] = h.in ? ldo(vy, q) : 0.0f;
        qz[4] = h.in ? ldo(vz, q) : 0.0f;
        const uint32_t pz = (uint32_t)z * sz;
        const HaloV hx = el_halo_load(vx, h, pz), hy = el_halo_load(vy, h, pz),
                    hz = el_halo_load(vz, h, pz);
        __syncthreads();  // the previous plane's reads are done
        el_stage(sv[0], h, qx[2], hx);
        el_stage(sv[1], h, qy[2], hy);
        el_stage(sv[2], h, qz[2], hz);
        __syncthreads();
        if (upd) {
            const float exx = fdx(sv[0], h, i0);
            const float eyy = fdy(sv[1], h, i1);
            const float ezz = fdq(qz, i2);
            const float dxy = fdy(sv[0], h, i1) + fdx(sv[1], h, i0);
            const float dxz = fdq(qx, i2) + fdx(sv[2], h, i0);
            const float dyz = fdq(qy, i2) + fdy(sv[2], h, i1);
            const float l = ldo(lam, c), m = ldo(mu, c);
            const float pxx = ldo(G(d.adj[3] + boff), c), pyy = ldo(G(d.adj[4] + boff), c),
                        pzz = ldo(G(d.adj[5] + boff), c);
            const float trs = ((pxx + pyy) + pzz) / (3.0f * l + 2.0f * m);
            const float im2 = 0.5f / m;
            const float hxx = (pxx - l * trs) * im2, hyy = (pyy - l * trs) * im2,
                        hzz = (pzz - l * trs) * im2;
            const float hxy = ldo(G(d.adj[6] + boff), c) / m,
                        hxz = ldo(G(d.adj[7] + boff), c) / m,
                        hyz = ldo(G(d.adj[8] + boff), c) / m;
            gfloat* __restrict__ gl = G(d.grad[0] + boff);
            gfloat* __restrict__ gm = G(d.grad[1] + boff);
            sto(gl, c, ldo(gl, c) + dt * (trs * ((exx + eyy) + ezz)));
            sto(gm, c, ldo(gm, c) + dt * (2.0f * (hxx * exx + hyy * eyy + hzz * ezz) +
                                          hxy * dxy + hxz * dxz + hyz * dyz));
            if (REV) {
                const float l2m = l + 2.0f * m;
                gfloat* __restrict__ sg = G(d.f[F_SXX] + boff);
                sto(sg, c, ldo(sg, c) - dt * (l2m * exx + l * (eyy + ezz)));
                sg = G(d.f[F_SYY] + boff);
                sto(sg, c, ldo(sg, c) - dt * (l2m * eyy + l * (exx + ezz)));
                sg = G(d.f[F_SZZ] + boff);
                sto(sg, c, ldo(sg, c) - dt * (l2m * ezz + l * (exx + eyy)));
                sg = G(d.f[F_SXY] + boff);
                sto(sg, c, ldo(sg, c) - dt * (m * dxy));
                sg = G(d.f[F_SXZ] + boff);
                sto(sg, c, ldo(sg, c) - dt * (m * dxz));
                sg = G(d.f[F_SYZ] + boff);
                sto(sg, c, ldo(sg, c) - dt * (m * dyz));
            }
        }
        qshift(qx);
        qshift(qy);
        qshift(qz);
    }
}

// Point terms of the reverse of step tcount[shot] (blockIdx.y = shot).
// phase 0: receivers feed the adjoint, vh += res, i.e. w += b res;
// phase 1: sources feed g_b += dt (w/b) f(t) and, REV, leave the state
// (v -= dt b f(t)).
// Each point is applied in the block that owns it; the exchange that follows
// refreshes the halo copies elsewhere.
__device__ __forceinline__ long long el_owner_cell(const KGeo& g, const int* p)
{
    int bc[3], lc[3];
    for (int a = 0; a < 3; ++a) {
        bc[a] = p[a] / g.B[a];
        lc[a] = p[a] - bc[a] * g.B[a] + g.h;
    }
    const int b = (bc[2] * g.nb[1] + bc[1]) * g.nb[0] + bc[0];
    return cell_index(g, b, lc[0], lc[1], lc[2]);
}

__global__ void el_adj_point_kernel(const ElDev* __restrict__ tab, const int* __restrict__ tcount,
                                    int phase, int rev)
{
    const ElDev& d = tab[blockIdx.y];
    const SrcRec& sr = d.sr;
    const int t = tcount[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const KGeo& g = d.g;
    if (phase == 0) {
        if (i >= sr.nrec || t < 0 || t >= sr.rec_nt) return;
        const long long c = el_owner_cell(g, sr.rec + 3 * i);
        d.adj[sr.rec_comp][c] += d.f[F_B][c] * d.res[(long long)t * sr.nrec + i];
        return;
    }
    if (i >= sr.nsrc || t < 0 || t >= sr.src_nt) return;
    const int* p = sr.src + 4 * i;
    const long long c = el_owner_cell(g, p);
    const float wv = sr.wav[(long long)t * sr.nsrc + i];
    const float bc = d.f[F_B][c];
    d.grad[2][c] += d.dt * ((d.adj[p[3]][c] / bc) * wv);
    if (rev) {
        float* v = d.f[p[3]];
        v[c] = v[c] - (d.dt * bc) * wv;
    }
}

// (w, p) -> (vh, sh) = (w / b, C^-1 p) on the owned cells, for returning the
// adjoint state itself
__global__ __launch_bounds__(EL_WG) void el_adj_unscale_kernel(const ElDev* __restrict__ tab, int nshot)
{
    const Tile T = decode_tile(tab, nshot);
    const ElDev& d = tab[T.shot];
    if (T.b >= d.g.nblk) return;  // padding tile
    const KGeo& g = d.g;
    const int h = g.h, x = T.x, y = T.y;
    if (x < h || x >= h + g.B[0] || y < h || y >= h + g.B[1]) return;
    const int z0 = max(T.zs, h), z1 = min(T.ze, h + g.B[2]) - 1;
    for (int z = z0; z <= z1; ++z) {
        const long long c = (long long)T.b * g.bstride + g.xoff + z * g.slice + (long long)y * g.pitch + x;
        const float l = d.f[F_LAM][c], m = d.f[F_MU][c], bc = d.f[F_B][c];
        for (int k = 0; k < 3; ++k) d.adj[k][c] = d.adj[k][c] / bc;
        const float pxx = d.adj[3][c], pyy = d.adj[4][c], pzz = d.adj[5][c];
        const float trs = ((pxx + pyy) + pzz) / (3.0f * l + 2.0f * m);
        const float im2 = 0.5f / m;
        d.adj[3][c] = (pxx - l * trs) * im2;
        d.adj[4][c] = (pyy - l * trs) * im2;
        d.adj[5][c] = (pzz - l * trs) * im2;
        for (int k = 6; k < 9; ++k) d.adj[k][c] = d.adj[k][c] / m;
    }
}

// Checkpoint slot j-1 <-> state (owned cells of the 9 fields).  save: after
// the forward reached step t = j*ckpt < nsteps; restore: before reversing
// step t with t+1 = j*ckpt < nsteps.  Other steps leave everything alone.
__global__ __launch_bounds__(EL_WG) void el_ckpt_kernel(const ElDev* __restrict__ tab, int nshot,
                                                      const int* __restrict__ tcount, int ckpt,
                                                      int nsteps, int restore)
{
    const Tile T = decode_tile(tab, nshot);
    const ElDev& d = tab[T.shot];
    if (T.b >= d.g.nblk) return;  // padding tile
    const KGeo& g = d.g;
    const int ts = tcount[T.shot] + (restore ? 1 : 0);
    if (ts % ckpt != 0 || ts < ckpt || ts >= nsteps) return;
    const int h = g.h, x = T.x, y = T.y;
    if (x < h || x >= h + g.B[0] || y < h || y >= h + g.B[1]) return;
    float* slot = d.cpt + (long long)(ts / ckpt - 1) * 9 * d.fstride;
    const int z0 = max(T.zs, h), z1 = min(T.ze, h + g.B[2]) - 1;
    const long long boff = (long long)T.b * g.bstride + g.xoff;
    for (int k = 0; k < 9; ++k) {
        float* st = d.f[k];
        float* cp = slot + (long long)k * d.fstride;
        for (int z = z0; z <= z1; ++z) {
            const long long c = boff + z * g.slice + (long long)y * g.pitch + x;
            if (restore) st[c] = cp[c];
            else cp[c] = st[c];
        }
    }
}

// res = rec - obs, and per-workgroup FP64 partial sums of res^2 (wave64 shuffles)
__global__ __launch_bounds__(256) void el_residual_kernel(const float* __restrict__ rec,
                                                          const float* __restrict__ obs,
                                                          float* __restrict__ res, long long n,
                                                          double* __restrict__ part)
{
    double acc = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const float r = rec[i] - obs[i];
        res[i] = r;
        acc += (double)r * (double)r;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    __shared__ double ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

__global__ void el_sum_kernel(const double* __restrict__ part, int n, double* __restrict__ out)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 64) acc += part[i];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (threadIdx.x == 0) *out = acc;
}

// ---------------------------------------------------------------- host side
// Tiling of the stage kernels (fused = false: 64 x 4 tiles over the stress
// boxes) or of the fused step kernel (60 x 12 velocity tiles over the
// velocity boxes).
static void el_tiles(ElDev& d, int zc_want, bool fused = false)
{
    const KGeo& g = d.g;
    // Stage update boxes lie in [2, S-2), velocity boxes in [4, S-4); on an
    // axis with one block and both global faces frozen they are the owned
    // cells [h, h+B) at every stage.
    for (int a = 0; a < 3; ++a) {
        const bool owned = g.nb[a] == 1 && !((g.periodic >> a) & 1) && ((g.fmask >> (2 * a)) & 3) == 3;
        const int o = fused ? 4 : 2;
        d.o[a] = owned ? g.h : o;
        d.w[a] = owned ? g.B[a] : g.S[a] - 2 * o;
    }
    d.ntx = ceil_div(d.w[0], fused ? EF_IX : EL_TX);
    d.nty = ceil_div(d.w[1], fused ? EF_IY : EL_TY);
    const int nz = d.w[2];
    long long tiles = (long long)d.ntx * d.nty * g.nblk;
    int zc = zc_want;
    if (zc <= 0) {  // automatic: about 2M threads, at least 8 (fused: 16) planes each
        const long long want = fused ? 2048 : 8192;
        const int nzc = (int)((want + tiles - 1) / tiles);
        zc = ceil_div(nz, nzc < 1 ? 1 : nzc);
        if (zc < (fused ? 16 : 8)) zc = fused ? 16 : 8;
    }
    if (zc > nz) zc = nz;
    d.zc = zc;
    d.nzc = ceil_div(nz, zc);
    d.ntiles = d.ntx * d.nty * d.nzc * g.nblk;
    d.ngrid = (d.ntiles + 7) & ~7;
}

static SrcRec el_srcrec(const b2r_el_shot* s)
{
    SrcRec r;
    r.src = s->d_src;
    r.wav = s->d_wav;
    r.nsrc = s->nsrc;
    r.src_nt = s->src_nt;
    r.rec = s->d_rec;
    r.recv = s->d_recv;
    r.nrec = s->nrec;
    r.rec_comp = s->rec_comp;
    r.rec_nt = s->rec_nt;
    return r;
}

extern "C" int b2r_el_create(b2r_ctx* ctx, float dt, const float inv_h[3], b2r_el_shot** out)
{
    B2R_CHECK(ctx && inv_h && out, B2R_E_ARG, "null argument");
    B2R_CHECK(ctx->g.elem_size == 4, B2R_E_UNSUPPORTED, "elastic fields are f32");
    B2R_CHECK(ctx->g.radius == 4, B2R_E_ARG,
              "elastic needs radius 4 (two radius-2 stages per step), got %d", ctx->g.radius);
    B2R_CHECK(ctx->k.bstride * 4 < (1ll << 32), B2R_E_UNSUPPORTED,
              "elastic blocks are limited to 4 GiB per field (32-bit offsets)");
    if (int rc = b2r_activate(ctx)) return rc;
    b2r_el_shot* s = new b2r_el_shot();
    s->ctx = ctx;
    s->t = 0;
    memset(&s->d, 0, sizeof(s->d));
    s->d.g = ctx->k;
    s->d.dt = dt;
    for (int a = 0; a < 3; ++a) s->d.ih[a] = inv_h[a];
    for (int k = 0; k < F_NUM; ++k) {
        if (int rc = b2r_field_alloc(ctx, &s->fields[k])) {
            for (int j = 0; j < k; ++j) (void)b2r_field_free(ctx, s->fields[j]);
            delete s;
            return rc;
        }
        s->d.f[k] = (float*)s->fields[k];
    }
    el_tiles(s->d, ctx->tune.el_zc);
    *out = s;
    return B2R_OK;
}

extern "C" int b2r_el_destroy(b2r_el_shot* s)
{
    if (!s) return B2R_OK;
    b2r_ctx* c = s->ctx;
    (void)hipStreamSynchronize(c->stream);
    for (int k = 0; k < F_NUM; ++k) b2r_field_free(c, s->fields[k]);
    for (int k = 0; k < 9; ++k)
        if (s->pp[k]) b2r_field_free(c, s->pp[k]);
    if (s->d.disp) b2r_field_free(c, s->d.disp);
    void* ptrs[] = {s->d_src, s->d_wav, s->d_rec, s->d_recv};
    for (void* p : ptrs)
        if (p) b2r_dev_free(c, p);
    delete s;
    return B2R_OK;
}

extern "C" int b2r_el_field(const b2r_el_shot* s, int which, void** d_field)
{
    B2R_CHECK(s && d_field && which >= 0 && which < F_NUM, B2R_E_ARG, "bad argument");
    *d_field = s->fields[which];
    return B2R_OK;
}

extern "C" int b2r_el_set_medium(b2r_el_shot* s, const float* lam, const float* mu, const float* b)
{
    B2R_CHECK(s && lam && mu && b, B2R_E_ARG, "null argument");
    b2r_ctx* c = s->ctx;
    if (int rc = b2r_field_upload(c, s->fields[F_LAM], lam)) return rc;
    if (int rc = b2r_field_upload(c, s->fields[F_MU], mu)) return rc;
    if (int rc = b2r_field_upload(c, s->fields[F_B], b)) return rc;
    // halos of the medium once: it never changes
    void* m[3] = {s->fields[F_LAM], s->fields[F_MU], s->fields[F_B]};
    if (int rc = b2r_direct_exchange(c, m, 3)) return rc;
    B2R_HIP(hipStreamSynchronize(c->stream));
    return B2R_OK;
}

extern "C" int b2r_el_reset(b2r_el_shot* s)
{
    B2R_CHECK(s, B2R_E_ARG, "null shot");
    b2r_ctx* c = s->ctx;
    for (int k = 0; k < 9; ++k)
        B2R_HIP(hipMemsetAsync(s->fields[k], 0, (size_t)c->L.field_elems * 4, c->stream));
    if (s->d.disp) B2R_HIP(hipMemsetAsync(s->d.disp, 0, (size_t)c->L.field_elems * 4, c->stream));
    s->t = 0;
    return B2R_OK;
}

extern "C" int b2r_el_enable_displacement(b2r_el_shot* s)
{
    B2R_CHECK(s, B2R_E_ARG, "null shot");
    if (!s->d.disp) {
        void* p = nullptr;
        if (int rc = b2r_field_alloc(s->ctx, &p)) return rc;
        s->d.disp = (float*)p;
    }
    return B2R_OK;
}

static int upload_buf(b2r_ctx* c, void** dst, const void* src, size_t bytes)
{
    if (*dst) {
        if (int rc = b2r_dev_free(c, *dst)) return rc;
        *dst = nullptr;
    }
    if (bytes == 0) return B2R_OK;
    if (int rc = b2r_dev_alloc(c, bytes, dst)) return rc;
    B2R_HIP(hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    B2R_HIP(hipStreamSynchronize(c->stream));
    return B2R_OK;
}

extern "C" int b2r_el_set_sources(b2r_el_shot* s, int nsrc, const int* xyz, const int* comp, int nt,
                                  const float* wav)
{
    B2R_CHECK(s && nsrc >= 0 && nt >= 0 && (nsrc == 0 || (xyz && comp && wav)), B2R_E_ARG,
              "bad argument");
    const b2r_geom& g = s->ctx->g;
    std::vector<int> t(4 * (size_t)nsrc);
    for (int i = 0; i < nsrc; ++i) {
        B2R_CHECK(xyz[3 * i] >= 0 && xyz[3 * i] < g.env_x && xyz[3 * i + 1] >= 0 &&
                  xyz[3 * i + 1] < g.env_y && xyz[3 * i + 2] >= 0 && xyz[3 * i + 2] < g.env_z,
                  B2R_E_ARG, "source %d outside the grid", i);
        B2R_CHECK(comp[i] >= 0 && comp[i] <= 2, B2R_E_ARG, "source component must be 0..2");
        for (int a = 0; a < 3; ++a) t[4 * i + a] = xyz[3 * i + a];
        t[4 * i + 3] = comp[i];
    }
    if (int rc = upload_buf(s->ctx, (void**)&s->d_src, t.data(), t.size() * sizeof(int))) return rc;
    if (int rc = upload_buf(s->ctx, (void**)&s->d_wav, wav, (size_t)nt * nsrc * sizeof(float)))
        return rc;
    s->nsrc = nsrc;
    s->src_nt = nt;
    return B2R_OK;
}

extern "C" int b2r_el_set_receivers(b2r_el_shot* s, int nrec, const int* xyz, int comp, int nt)
{
    B2R_CHECK(s && nrec >= 0 && nt >= 0 && comp >= 0 && comp < 9 && (nrec == 0 || xyz),
              B2R_E_ARG, "bad argument");
    const b2r_geom& g = s->ctx->g;
    for (int i = 0; i < nrec; ++i)
        B2R_CHECK(xyz[3 * i] >= 0 && xyz[3 * i] < g.env_x && xyz[3 * i + 1] >= 0 &&
                  xyz[3 * i + 1] < g.env_y && xyz[3 * i + 2] >= 0 && xyz[3 * i + 2] < g.env_z,
                  B2R_E_ARG, "receiver %d outside the grid", i);
    if (int rc = upload_buf(s->ctx, (void**)&s->d_rec, xyz, (size_t)nrec * 3 * sizeof(int)))
        return rc;
    if (s->d_recv) {
        if (int rc = b2r_dev_free(s->ctx, s->d_recv)) return rc;
        s->d_recv = nullptr;
    }
    if ((size_t)nrec * nt) {
        if (int rc = b2r_dev_alloc(s->ctx, (size_t)nrec * nt * sizeof(float), (void**)&s->d_recv))
            return rc;
        B2R_HIP(hipMemsetAsync(s->d_recv, 0, (size_t)nrec * nt * sizeof(float), s->ctx->stream));
    }
    s->nrec = nrec;
    s->rec_comp = comp;
    s->rec_nt = nt;
    return B2R_OK;
}

extern "C" int b2r_el_get_receivers(b2r_el_shot* s, float* h_rec)
{
    B2R_CHECK(s && h_rec, B2R_E_ARG, "null argument");
    if ((size_t)s->nrec * s->rec_nt == 0) return B2R_OK;
    B2R_HIP(hipMemcpyAsync(h_rec, s->d_recv, (size_t)s->nrec * s->rec_nt * sizeof(float),
                           hipMemcpyDeviceToHost, s->ctx->stream));
    B2R_HIP(hipStreamSynchronize(s->ctx->stream));
    return B2R_OK;
}

extern "C" int b2r_el_get_state(b2r_el_shot* s, int which, float* h_global)
{
    B2R_CHECK(s && h_global && which >= -1 && which < F_NUM, B2R_E_ARG, "bad argument");
    const void* f = which < 0 ? (const void*)s->d.disp : s->fields[which];
    B2R_CHECK(f, B2R_E_ARG, "displacement not enabled");
    return b2r_field_download(s->ctx, f, h_global);
}

extern "C" int b2r_el_set_state(b2r_el_shot* s, int which, const float* h_global)
{
    B2R_CHECK(s && h_global && which >= 0 && which < 9, B2R_E_ARG, "bad argument");
    return b2r_field_upload(s->ctx, s->fields[which], h_global);
}

// ---------------------------------------------------------------- stepping
struct ElBatch {
    std::vector<b2r_el_shot*> shots;
    ElDev* d_tab = nullptr;   // device copy of the shot table
    ElDev* d_tab1 = nullptr;  // fused: the table with the two state sets swapped
    bool fused = false;       // one el_step_kernel launch per step (ping-pong state)
    int par = 0;              // fused: the state is in the second set (pp)
    int* d_t = nullptr;       // per shot step counters (device)
    int total_tiles = 0;      // workgroups of a batched launch
    int max_src = 0;          // max over shots of sources x blocks
    int max_rec = 0;          // max over shots of receivers
    hipStream_t stream;
};

// One step (inner step i of the B2R window) of every shot in the batch.
// plain: per shot, a stress launch, separate vx, vy, vz launches and the
// point terms (eqwp_gpu_main's structure, eqwp.cu:679-684); batched: one
// stress and one three-component velocity launch over the table of shots,
// one launch each for the point terms of all shots.
static int el_step_launch(ElBatch& bt, int i, bool plain)
{
    const int nshot = (int)bt.shots.size();
    dim3 blk(EL_TX, EL_TY);
    if (bt.fused) {
        // read the set holding the state, write the other, then apply the
        // point terms to the new state through the swapped table
        const ElDev* cur = bt.par ? bt.d_tab1 : bt.d_tab;
        const ElDev* nxt = bt.par ? bt.d_tab : bt.d_tab1;
        el_step_kernel<<<bt.total_tiles, dim3(EF_TX, EF_TY), 0, bt.stream>>>(cur, nshot, i);
        if (bt.max_src)
            el_inject_kernel<<<dim3(ceil_div(bt.max_src, 64), nshot), 64, 0, bt.stream>>>(
                nxt, 2 * i + 1, bt.d_t);
        if (bt.max_rec)
            el_record_kernel<<<dim3(ceil_div(bt.max_rec, 64), nshot), 64, 0, bt.stream>>>(nxt,
                                                                                         bt.d_t);
        bt.par ^= 1;
    } else if (plain) {
        for (int k = 0; k < nshot; ++k) {
            // a one-entry table (tile0 = 0 for plain tables)
            const ElDev* tk = bt.d_tab + k;
            const b2r_el_shot* s = bt.shots[k];
            const int nt = s->d.ngrid;
            el_stress_kernel<<<nt, blk, 0, bt.stream>>>(tk, 1, 2 * i);
            el_velocity_kernel<1><<<nt, blk, 0, bt.stream>>>(tk, 1, 2 * i + 1);
            el_velocity_kernel<2><<<nt, blk, 0, bt.stream>>>(tk, 1, 2 * i + 1);
            el_velocity_kernel<4><<<nt, blk, 0, bt.stream>>>(tk, 1, 2 * i + 1);
            if (s->nsrc)
                el_inject_kernel<<<dim3(ceil_div(s->nsrc * s->d.g.nblk, 64), 1), 64, 0, bt.stream>>>(
                    tk, 2 * i + 1, bt.d_t + k);
            if (s->nrec)
                el_record_kernel<<<dim3(ceil_div(s->nrec, 64), 1), 64, 0, bt.stream>>>(tk, bt.d_t + k);
        }
    } else {
        el_stress_kernel<<<bt.total_tiles, blk, 0, bt.stream>>>(bt.d_tab, nshot, 2 * i);
        el_velocity_kernel<7><<<bt.total_tiles, blk, 0, bt.stream>>>(bt.d_tab, nshot, 2 * i + 1);
        if (bt.max_src)
            el_inject_kernel<<<dim3(ceil_div(bt.max_src, 64), nshot), 64, 0, bt.stream>>>(
                bt.d_tab, 2 * i + 1, bt.d_t);
        if (bt.max_rec)
            el_record_kernel<<<dim3(ceil_div(bt.max_rec, 64), nshot), 64, 0, bt.stream>>>(bt.d_tab,
                                                                                         bt.d_t);
    }
    el_tick_kernel<<<1, 64, 0, bt.stream>>>(bt.d_t, nshot, 1);
    B2R_LAUNCH_CHECK();
    return B2R_OK;
}

static int el_exchange(ElBatch& bt)
{
    for (b2r_el_shot* s : bt.shots) {
        const KGeo& g = s->ctx->k;
        if (g.nblk == 1 && !g.periodic) continue;  // nothing to exchange
        void* f[9];
        for (int k = 0; k < 9; ++k) f[k] = bt.par ? s->pp[k] : s->fields[k];
        // the exchange runs on the batch stream
        hipStream_t keep = s->ctx->stream;
        s->ctx->stream = bt.stream;
        int rc = b2r_direct_exchange(s->ctx, f, 9);
        s->ctx->stream = keep;
        if (rc) return rc;
    }
    return B2R_OK;
}

// Uploads a shot table followed by its first-tile array (see decode_tile).
static int el_upload_table(const std::vector<ElDev>& tab, hipStream_t st, ElDev** d_tab)
{
    const size_t n = tab.size();
    std::vector<int> t0(n);
    for (size_t k = 0; k < n; ++k) t0[k] = tab[k].tile0;
    const size_t bytes = sizeof(ElDev) * n + sizeof(int) * n;
    B2R_CHECK(hipMalloc(d_tab, bytes) == hipSuccess, B2R_E_NOMEM, "table allocation failed");
    B2R_HIP(hipMemcpyAsync(*d_tab, tab.data(), sizeof(ElDev) * n, hipMemcpyHostToDevice, st));
    B2R_HIP(hipMemcpyAsync(*d_tab + n, t0.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    return B2R_OK;
}

// Whether a batched step runs the fused kernel: el_kernel 1, or 0 where no
// global face is open (an open face keeps cells outside the velocity box in
// the state, which the fused step does not carry over).
static int el_fused_choice(b2r_el_shot* const* shots, int nshots, bool* fused)
{
    *fused = true;
    for (int k = 0; k < nshots; ++k) {
        const b2r_ctx* c = shots[k]->ctx;
        const bool ok = closed_faces(c->k);
        B2R_CHECK(ok || c->tune.el_kernel != 1, B2R_E_UNSUPPORTED,
                  "el_kernel 1 needs every global face frozen or periodic");
        if (!ok || c->tune.el_kernel == 2) *fused = false;
    }
    return B2R_OK;
}

// Device table of the shots (their ElDev with sources and receivers) and
// step counters.  Batched tables number the workgroups of all shots
// consecutively (tile0); plain tables give every shot tile0 = 0.  A fused
// batch also gets each shot's second state set and the table with the sets
// swapped.
static int el_batch_init(ElBatch& bt, b2r_el_shot* const* shots, int nshots, bool plain,
                         bool fused = false)
{
    B2R_CHECK(shots && nshots >= 1 && nshots <= EL_MAX_BATCH, B2R_E_ARG, "1..%d shots",
              EL_MAX_BATCH);
    bt.shots.assign(shots, shots + nshots);
    for (int k = 0; k < nshots; ++k) B2R_CHECK(shots[k], B2R_E_ARG, "null shot %d", k);
    const int R = shots[0]->ctx->g.R;
    std::vector<ElDev> tab(nshots);
    std::vector<int> t0(nshots);
    int tiles = 0;
    for (int k = 0; k < nshots; ++k) {
        b2r_el_shot* s = shots[k];
        B2R_CHECK(s->ctx->device == shots[0]->ctx->device, B2R_E_ARG,
                  "all shots of a batch must be on one device");
        B2R_CHECK(s->ctx->g.R == R, B2R_E_ARG, "all shots of a batch must share R");
        s->d.g = s->ctx->k;
        const bool fz = fused && !plain;
        el_tiles(s->d, fz ? s->ctx->tune.el_step_zc : s->ctx->tune.el_zc, fz);
        s->d.sr = el_srcrec(s);
        for (int j = 0; j < 9; ++j) s->d.nx[j] = nullptr;
        tab[k] = s->d;
        tab[k].tile0 = plain ? 0 : tiles;
        tiles += s->d.ngrid;
        bt.max_src = std::max(bt.max_src, s->nsrc * s->d.g.nblk);
        bt.max_rec = std::max(bt.max_rec, s->nrec);
        t0[k] = s->t;
    }
    bt.total_tiles = plain ? 0 : tiles;
    bt.stream = shots[0]->ctx->stream;
    // work queued on the shots' own context streams comes first
    for (int k = 0; k < nshots; ++k)
        if (shots[k]->ctx->stream != bt.stream) B2R_HIP(hipStreamSynchronize(shots[k]->ctx->stream));
    bt.fused = fused && !plain;
    bt.par = 0;
    if (bt.fused) {
        std::vector<ElDev> tab1(tab);
        for (int k = 0; k < nshots; ++k) {
            b2r_el_shot* s = shots[k];
            for (int j = 0; j < 9; ++j) {
                if (!s->pp[j])
                    if (int rc = b2r_field_alloc(s->ctx, &s->pp[j])) return rc;
                tab[k].nx[j] = (float*)s->pp[j];
                tab1[k].f[j] = (float*)s->pp[j];
                tab1[k].nx[j] = (float*)s->fields[j];
            }
        }
        if (int rc = el_upload_table(tab1, bt.stream, &bt.d_tab1)) return rc;
    }
    if (int rc = el_upload_table(tab, bt.stream, &bt.d_tab)) return rc;
    B2R_HIP(hipMalloc(&bt.d_t, sizeof(int) * nshots));
    B2R_HIP(hipMemcpyAsync(bt.d_t, t0.data(), sizeof(int) * nshots, hipMemcpyHostToDevice,
                           bt.stream));
    return B2R_OK;
}

// Copies the state between the two sets of a fused batch: to_second = true
// before the first step (every cell the steps do not write, the frozen frame
// among them, must be alike in both), false after a run that ended in the
// second set.
static int el_fused_copy(ElBatch& bt, bool to_second)
{
    for (b2r_el_shot* s : bt.shots) {
        const size_t bytes = (size_t)s->ctx->L.field_elems * 4;
        for (int j = 0; j < 9; ++j)
            B2R_HIP(hipMemcpyAsync(to_second ? s->pp[j] : s->fields[j],
                                   to_second ? s->fields[j] : s->pp[j], bytes,
                                   hipMemcpyDeviceToDevice, bt.stream));
    }
    return B2R_OK;
}

static void el_batch_free(ElBatch& bt)
{
    (void)hipStreamSynchronize(bt.stream);
    if (bt.d_tab) (void)hipFree(bt.d_tab);
    if (bt.d_tab1) (void)hipFree(bt.d_tab1);
    if (bt.d_t) (void)hipFree(bt.d_t);
    bt.d_tab1 = nullptr;
    bt.d_tab = nullptr;
    bt.d_t = nullptr;
}

// steps of one R-window group: exchange, then n inner steps
static int el_group(ElBatch& bt, int n, bool plain)
{
    if (int rc = el_exchange(bt)) return rc;
    for (int i = 0; i < n; ++i) {
        if (int rc = el_step_launch(bt, i, plain)) return rc;
        if (plain) B2R_HIP(hipStreamSynchronize(bt.stream));  // eqwp.cu:685
    }
    return B2R_OK;
}

extern "C" int b2r_el_forward_batch(b2r_el_shot* const* shots, int nshots, int nsteps, int mode,
                                    float* ms)
{
    B2R_CHECK(nsteps >= 0 && mode >= 0 && mode <= 2, B2R_E_ARG, "bad nsteps or mode");
    const bool plain = mode == 0, graph = mode == 2;
    B2R_CHECK(shots && nshots >= 1, B2R_E_ARG, "no shots");
    for (int k = 0; k < nshots; ++k) B2R_CHECK(shots[k], B2R_E_ARG, "null shot %d", k);
    bool fused = false;
    if (!plain)
        if (int rc = el_fused_choice(shots, nshots, &fused)) return rc;
    ElBatch bt;
    if (int rc = el_batch_init(bt, shots, nshots, plain, fused)) {
        el_batch_free(bt);
        return rc;
    }
    b2r_ctx* c0 = shots[0]->ctx;
    if (int rc = b2r_activate(c0)) return rc;
    const int R = c0->g.R;
    int rc = B2R_OK;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph_ = nullptr;
    // graphs replay one chunk of whole R-windows; the step counters live on
    // the device so one graph serves every chunk
    int chunk = R;
    while (chunk < 32 && chunk * 2 <= nsteps) chunk *= 2;
    chunk -= chunk % R;
    if (chunk < R) chunk = R;
    // a fused graph must end in the set it starts from: an even step count
    if (graph && nsteps >= chunk && !(bt.fused && (chunk & 1))) {
        if (hipStreamBeginCapture(bt.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            b2r_set_error("hipStreamBeginCapture failed");
            rc = B2R_E_HIP;
        }
        for (int s = 0; s < chunk && !rc; s += R) rc = el_group(bt, R, false);
        hipError_t e = hipStreamEndCapture(bt.stream, &graph_);
        if (!rc && e != hipSuccess) {
            b2r_set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
            rc = B2R_E_HIP;
        }
        if (!rc && hipGraphInstantiate(&exec, graph_, nullptr, nullptr, 0) != hipSuccess) {
            b2r_set_error("hipGraphInstantiate failed");
            rc = B2R_E_HIP;
        }
    }
    if (!rc) rc = b2r_time_begin(c0, ms);
    if (!rc && bt.fused) rc = el_fused_copy(bt, true);
    int done = 0;
    if (!rc && exec) {
        for (; done + chunk <= nsteps && !rc; done += chunk)
            if (hipGraphLaunch(exec, bt.stream) != hipSuccess) {
                b2r_set_error("hipGraphLaunch failed");
                rc = B2R_E_HIP;
            }
    }
    while (!rc && done < nsteps) {
        const int n = nsteps - done < R ? nsteps - done : R;
        rc = el_group(bt, n, plain);
        done += n;
    }
    // a fused run that ended in the second set copies the state back
    if (!rc && bt.par) rc = el_fused_copy(bt, false);
    if (!rc) rc = b2r_time_end(c0, ms);
    if (!rc) B2R_HIP(hipStreamSynchronize(bt.stream));
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (!rc)
        for (int k = 0; k < nshots; ++k) shots[k]->t += nsteps;
    el_batch_free(bt);
    return rc;
}

extern "C" int b2r_el_forward(b2r_el_shot* s, int nsteps, int mode, float* ms)
{
    return b2r_el_forward_batch(&s, 1, nsteps, mode, ms);
}

// Device time of `steps` batched steps (mode 1) without the fixed cost of a
// run: a fused run copies the state into the second set first, once per call,
// which a tuning run of a few steps would otherwise charge to every step.
static int el_time_batch(b2r_el_shot* const* shots, int nshots, int steps, float* ms)
{
    float m0 = 0.0f, m1 = 0.0f;
    if (int rc = b2r_el_forward_batch(shots, nshots, 0, 1, &m0)) return rc;
    if (int rc = b2r_el_forward_batch(shots, nshots, steps, 1, &m1)) return rc;
    *ms = m1 > m0 ? m1 - m0 : 0.0f;
    return B2R_OK;
}

// Autotuner of the elastic stencil kernels (no counterpart in the reference):
// times the batched step of this shot (the fused step kernel where the faces
// allow it and the separate stress and three-component velocity kernels,
// each with the point terms) for each z-chunk candidate, keeps the fastest in
// the shot's context, appends one record line as b2r_autotune does, and
// leaves the shot reset (zero state, t = 0).
extern "C" int b2r_el_autotune(b2r_el_shot* s, const char* record_path, b2r_tuning* out)
{
    B2R_CHECK(s, B2R_E_ARG, "null shot");
    b2r_ctx* c = s->ctx;
    const b2r_tuning base = c->tune;
    const int zcs[] = {0, 16, 24, 32, 48, 64, 96, 128, 256};
    const bool can_fuse = closed_faces(c->k);
    const int warm = 2, iters = 6;
    // best z chunk and time per kernel (index 1 fused, 2 separate)
    float kms[3] = {1e30f, 1e30f, 1e30f};
    int kzc[3] = {0, base.el_step_zc, base.el_zc}, tried = 0;
    for (int kern = can_fuse ? 1 : 2; kern <= 2; ++kern)
        for (int zc : zcs) {
            c->tune = base;
            c->tune.el_kernel = kern;
            (kern == 1 ? c->tune.el_step_zc : c->tune.el_zc) = zc;
            float ms = 0;
            int rc = b2r_el_forward(s, warm, 1, nullptr);
            if (!rc) rc = el_time_batch(&s, 1, iters, &ms);
            if (rc) {
                c->tune = base;
                (void)b2r_el_reset(s);
                return rc;
            }
            ++tried;
            if (ms < kms[kern]) {
                kms[kern] = ms;
                kzc[kern] = zc;
            }
        }
    const int best_k = kms[1] < kms[2] ? 1 : 2;
    const float best_ms = kms[best_k];
    c->tune = base;
    c->tune.el_kernel = best_k;
    c->tune.el_step_zc = kzc[1];
    c->tune.el_zc = kzc[2];
    el_tiles(s->d, kzc[2]);
    if (int rc = b2r_el_reset(s)) return rc;
    if (out) *out = c->tune;
    if (record_path && *record_path) {
        FILE* f = fopen(record_path, "a");
        B2R_CHECK(f, B2R_E_ARG, "cannot open %s", record_path);
        fprintf(f, "kind=el_step S=%dx%dx%d blocks=%d R=%d radius=%d "
                   "best=el_kernel:%d,el_zc:%d,el_step_zc:%d ms_per_step=%.5f",
                c->L.S[0], c->L.S[1], c->L.S[2], c->L.nblocks, c->g.R, c->g.radius, best_k,
                kzc[2], kzc[1], best_ms / iters);
        if (can_fuse)
            fprintf(f, " fused_ms_per_step=%.5f separate_ms_per_step=%.5f", kms[1] / iters,
                    kms[2] / iters);
        fprintf(f, " tried=%d\n", tried);
        fclose(f);
    }
    return B2R_OK;
}

extern "C" int b2r_el_autotune_batch(b2r_el_shot* const* shots, int nshots, const char* record_path)
{
    B2R_CHECK(shots && nshots >= 1 && nshots <= EL_MAX_BATCH, B2R_E_ARG, "1..%d shots",
              EL_MAX_BATCH);
    std::vector<b2r_tuning> base(nshots);
    bool can_fuse = true;
    long long cells = 0;
    for (int k = 0; k < nshots; ++k) {
        B2R_CHECK(shots[k], B2R_E_ARG, "null shot %d", k);
        const b2r_ctx* c = shots[k]->ctx;
        base[k] = c->tune;
        can_fuse = can_fuse && closed_faces(c->k);
        cells += (long long)c->g.env_x * c->g.env_y * c->g.env_z;
    }
    auto restore = [&]() {
        for (int k = 0; k < nshots; ++k) shots[k]->ctx->tune = base[k];
    };
    const int zcs[] = {0, 16, 24, 32, 48, 64, 96, 128};
    const int warm = 2, iters = 6;
    float kms[3] = {1e30f, 1e30f, 1e30f};
    int kzc[3] = {0, 0, 0}, tried = 0;
    for (int kern = can_fuse ? 1 : 2; kern <= 2; ++kern)
        for (int zc : zcs) {
            for (int k = 0; k < nshots; ++k) {
                b2r_tuning& t = shots[k]->ctx->tune;
                t = base[k];
                t.el_kernel = kern;
                (kern == 1 ? t.el_step_zc : t.el_zc) = zc;
            }
            float ms = 0;
            int rc = b2r_el_forward_batch(shots, nshots, warm, 1, nullptr);
            if (!rc) rc = el_time_batch(shots, nshots, iters, &ms);
            if (rc) {
                restore();
                for (int k = 0; k < nshots; ++k) (void)b2r_el_reset(shots[k]);
                return rc;
            }
            ++tried;
            if (ms < kms[kern]) {
                kms[kern] = ms;
                kzc[kern] = zc;
            }
        }
    const int best_k = kms[1] < kms[2] ? 1 : 2;
    for (int k = 0; k < nshots; ++k) {
        b2r_tuning& t = shots[k]->ctx->tune;
        t = base[k];
        t.el_kernel = best_k;
        if (can_fuse) t.el_step_zc = kzc[1];
        t.el_zc = kzc[2];
        if (int rc = b2r_el_reset(shots[k])) return rc;
    }
    if (record_path && *record_path) {
        FILE* f = fopen(record_path, "a");
        B2R_CHECK(f, B2R_E_ARG, "cannot open %s", record_path);
        fprintf(f, "kind=el_batch shots=%d cells=%lld best=el_kernel:%d,el_zc:%d,el_step_zc:%d "
                   "ms_per_step=%.5f",
                nshots, cells, best_k, kzc[2], kzc[1], kms[best_k] / iters);
        if (can_fuse)
            fprintf(f, " fused_ms_per_step=%.5f separate_ms_per_step=%.5f", kms[1] / iters,
                    kms[2] / iters);
        fprintf(f, " tried=%d\n", tried);
        fclose(f);
    }
    return B2R_OK;
}

// ---------------------------------------------------------------- gradient
extern "C" int b2r_el_misfit(b2r_el_shot* s, const float* h_obs, double* misfit)
{
    B2R_CHECK(s && h_obs && misfit, B2R_E_ARG, "null argument");
    b2r_ctx* c = s->ctx;
    const long long n = (long long)s->nrec * s->rec_nt;
    B2R_CHECK(n > 0, B2R_E_ARG, "no receivers");
    if (int rc = b2r_activate(c)) return rc;
    const int nb = 256;
    void* buf = nullptr;  // [obs n][res n][partials nb + 1]
    B2R_CHECK(hipMalloc(&buf, n * 8 + (nb + 1) * sizeof(double)) == hipSuccess, B2R_E_NOMEM,
              "misfit: device allocation failed");
    float* d_obs = (float*)buf;
    float* d_res = d_obs + n;
    double* d_part = (double*)(d_res + n + (n & 1));
    double ss = 0;
    hipError_t e = hipMemcpyAsync(d_obs, h_obs, n * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        el_residual_kernel<<<nb, 256, 0, c->stream>>>(s->d_recv, d_obs, d_res, n, d_part);
        el_sum_kernel<<<1, 64, 0, c->stream>>>(d_part, nb, d_part + nb);
        e = hipMemcpyAsync(&ss, d_part + nb, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(buf);
    B2R_CHECK(e == hipSuccess, B2R_E_HIP, "misfit: %s", hipGetErrorString(e));
    *misfit = 0.5 * ss;
    return B2R_OK;
}

static int copy_state(b2r_el_shot* s, hipStream_t st, float* const* dst, const float* const* src)
{
    const size_t bytes = (size_t)s->ctx->L.field_elems * 4;
    for (int k = 0; k < 9; ++k)
        B2R_HIP(hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyDeviceToDevice, st));
    return B2R_OK;
}

// Table entry of the adjoint propagation: the forward kernels on (w, p)
// with the step negated (see the adjoint kernels above).
static ElDev el_adj_entry(const ElDev& d)
{
    ElDev a = d;
    for (int k = 0; k < 9; ++k) a.f[k] = d.adj[k];
    a.dt = -d.dt;
    a.disp = nullptr;
    return a;
}

// The reverse sweep (rev) runs at R = 1 with every global face frozen or
// periodic, on any block decomposition; the plain per-shot loop on one block
// with frozen faces.
static bool el_grad_exchanges(const b2r_ctx* c) { return c->k.nblk > 1 || c->k.periodic; }

static int el_grad_check(b2r_el_shot* s, int nsteps, bool rev)
{
    B2R_CHECK(s, B2R_E_ARG, "null shot");
    b2r_ctx* c = s->ctx;
    B2R_CHECK(c->g.R == 1 && closed_faces(c->k), B2R_E_UNSUPPORTED,
              "the adjoint runs at R=1 with every global face frozen or periodic");
    B2R_CHECK(rev || (!el_grad_exchanges(c) && c->k.fmask == 0x3F), B2R_E_UNSUPPORTED,
              "the plain per-shot gradient runs on one block with frozen faces");
    B2R_CHECK(s->rec_nt >= nsteps && s->nrec > 0, B2R_E_ARG, "receivers must cover nsteps");
    B2R_CHECK(s->d.disp == nullptr, B2R_E_UNSUPPORTED, "disable the displacement output");
    return B2R_OK;
}

// Gradients out through the owned-cell copy: 3 grids, then optionally the 9
// fields of the adjoint state.
static int el_grad_out(b2r_el_shot* s, const float* grd, const float* adj, float* h_glam,
                       float* h_gmu, float* h_gb, float* h_adj0)
{
    b2r_ctx* c = s->ctx;
    const size_t fe = (size_t)c->L.field_elems;
    const long long gcells = (long long)c->g.env_x * c->g.env_y * c->g.env_z;
    void* f = nullptr;
    int rc = b2r_field_alloc(c, &f);
    for (int k = 0; k < 3 + (h_adj0 ? 9 : 0) && !rc; ++k) {
        float* out = k == 0 ? h_glam : (k == 1 ? h_gmu : (k == 2 ? h_gb : h_adj0 + (k - 3) * gcells));
        const float* srcp = k < 3 ? grd + k * fe : adj + (k - 3) * fe;
        if (hipMemcpyAsync(f, srcp, fe * 4, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
            b2r_set_error("gradient copy failed");
            rc = B2R_E_HIP;
        }
        if (!rc) rc = b2r_field_download(c, f, out);
    }
    if (f) b2r_field_free(c, f);
    return rc;
}

// residuals rec - obs on the device (res) and the FP64 misfit
static int el_residuals(b2r_el_shot* s, hipStream_t st, const float* h_obs, float* d_obs,
                        float* d_res, double* d_part, double* misfit)
{
    const long long n = (long long)s->nrec * s->rec_nt;
    const int nb = 256;
    B2R_HIP(hipMemcpyAsync(d_obs, h_obs, n * 4, hipMemcpyHostToDevice, st));
    el_residual_kernel<<<nb, 256, 0, st>>>(s->d_recv, d_obs, d_res, n, d_part);
    el_sum_kernel<<<1, 64, 0, st>>>(d_part, nb, d_part + nb);
    B2R_LAUNCH_CHECK();
    B2R_HIP(hipMemcpyAsync(misfit, d_part + nb, sizeof(double), hipMemcpyDeviceToHost, st));
    return B2R_OK;
}

// Plain per-shot gradient: the forward in the reference's structure (stress,
// vx, vy, vz launches and a synchronisation per step), states checkpointed
// every ckpt steps; every segment is recomputed with all its states kept on
// the device, then reversed with the adjoint kernels, one step at a time.
static int el_gradient_plain(b2r_el_shot* s, const float* h_obs, int nsteps, int ckpt, int flags,
                             double* misfit, float* h_glam, float* h_gmu, float* h_gb,
                             float* h_adj0)
{
    if (int rc = el_grad_check(s, nsteps, false)) return rc;
    b2r_ctx* c = s->ctx;
    if (int rc = b2r_activate(c)) return rc;
    const size_t fe = (size_t)c->L.field_elems, fb = fe * 4;
    const int ncp = (nsteps + ckpt - 1) / ckpt;
    const long long nobs = (long long)s->nrec * s->rec_nt;
    std::vector<void*> mem;
    auto alloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        mem.push_back(p);
        return p;
    };
    ElBatch bt;
    auto release = [&]() {
        el_batch_free(bt);
        (void)hipStreamSynchronize(c->stream);
        for (void* p : mem) (void)hipFree(p);
    };
    float* cp = (float*)alloc(fb * 9 * ncp);
    float* seg = (float*)alloc(fb * 9 * (ckpt + 1));
    float* adj = (float*)alloc(fb * 9);
    float* grd = (float*)alloc(fb * 3);
    float* obs = (float*)alloc(nobs * 8 + 8 + 257 * sizeof(double));
    // [state entry][adjoint entry][step counter]
    ElDev* d_tab = (ElDev*)alloc(2 * sizeof(ElDev) + 64);
    if (!cp || !seg || !adj || !grd || !obs || !d_tab) {
        release();
        b2r_set_error("gradient: device allocation of %zu MB failed",
                      fb * (9 * ncp + 9 * (ckpt + 1) + 12) >> 20);
        return B2R_E_NOMEM;
    }
    float* res = obs + nobs;
    double* part = (double*)(res + nobs + (nobs & 1));
    auto slot = [&](float* base, int j, float** out) {
        for (int k = 0; k < 9; ++k) out[k] = base + ((size_t)j * 9 + k) * fe;
    };
    int rc = B2R_OK;
    if (flags & 1) s->t = 0;
    else rc = b2r_el_reset(s);
    if (!rc) rc = el_batch_init(bt, &s, 1, true);
    // forward with checkpoints
    for (int j = 0; j < ncp && !rc; ++j) {
        float* dst[9];
        slot(cp, j, dst);
        rc = copy_state(s, c->stream, dst, s->d.f);
        const int n = std::min(ckpt, nsteps - j * ckpt);
        for (int i = 0; i < n && !rc; ++i) rc = el_group(bt, 1, true);
    }
    double ms_h = 0;
    if (!rc) rc = el_residuals(s, c->stream, h_obs, obs, res, part, &ms_h);
    if (!rc && hipMemsetAsync(adj, 0, fb * 9, c->stream) != hipSuccess) rc = B2R_E_HIP;
    if (!rc && hipMemsetAsync(grd, 0, fb * 3, c->stream) != hipSuccess) rc = B2R_E_HIP;
    if (!rc) B2R_HIP(hipStreamSynchronize(c->stream));
    ElDev d = s->d;
    d.g = c->k;
    for (int k = 0; k < 9; ++k) d.adj[k] = adj + k * fe;
    for (int k = 0; k < 3; ++k) d.grad[k] = grd + k * fe;
    ElDev* d_atab = d_tab + 1;
    int* d_tt = (int*)(d_tab + 2);
    d.sr = el_srcrec(s);
    d.res = res;
    ElDev ad = el_adj_entry(d);
    if (!rc && hipMemcpyAsync(d_atab, &ad, sizeof(ElDev), hipMemcpyHostToDevice, c->stream) !=
                   hipSuccess)
        rc = B2R_E_HIP;
    dim3 blk(EL_TX, EL_TY);
    for (int j = ncp - 1; j >= 0 && !rc; --j) {
        const int t0 = j * ckpt, t1 = std::min(nsteps, t0 + ckpt);
        // recompute the segment, keeping every state
        float* src[9];
        slot(cp, j, src);
        float* st[9];
        slot(seg, 0, st);
        rc = copy_state(s, c->stream, st, src);
        if (!rc) rc = copy_state(s, c->stream, s->d.f, st);
        if (!rc && hipMemcpyAsync(bt.d_t, &t0, sizeof(int), hipMemcpyHostToDevice, c->stream) !=
                       hipSuccess)
            rc = B2R_E_HIP;
        for (int t = t0; t < t1 && !rc; ++t) {
            rc = el_group(bt, 1, true);
            slot(seg, t - t0 + 1, st);
            if (!rc) rc = copy_state(s, c->stream, st, s->d.f);
        }
        for (int t = t1 - 1; t >= t0 && !rc; --t) {
            slot(seg, t - t0, d.fw0);
            slot(seg, t - t0 + 1, d.fw1);
            B2R_HIP(hipMemcpyAsync(d_tab, &d, sizeof(ElDev), hipMemcpyHostToDevice, c->stream));
            B2R_HIP(hipMemcpyAsync(d_tt, &t, sizeof(int), hipMemcpyHostToDevice, c->stream));
            el_adj_point_kernel<<<dim3(ceil_div(s->nrec, 64), 1), 64, 0, c->stream>>>(d_tab, d_tt, 0, 0);
            el_adj_vel_kernel<false><<<d.ngrid, blk, 0, c->stream>>>(d_tab, 1);
            if (s->nsrc)
                el_adj_point_kernel<<<dim3(ceil_div(s->nsrc, 64), 1), 64, 0, c->stream>>>(d_tab, d_tt, 1, 0);
            el_stress_kernel<<<d.ngrid, blk, 0, c->stream>>>(d_atab, 1, 0);
            el_adj_str_kernel<false><<<d.ngrid, blk, 0, c->stream>>>(d_tab, 1);
            el_velocity_kernel<1><<<d.ngrid, blk, 0, c->stream>>>(d_atab, 1, 1);
            el_velocity_kernel<2><<<d.ngrid, blk, 0, c->stream>>>(d_atab, 1, 1);
            el_velocity_kernel<4><<<d.ngrid, blk, 0, c->stream>>>(d_atab, 1, 1);
            // the table is rewritten next step: wait for these kernels first
            B2R_HIP(hipStreamSynchronize(c->stream));
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) {
                b2r_set_error("adjoint kernels: %s", hipGetErrorString(e));
                rc = B2R_E_HIP;
            }
        }
    }
    s->t = nsteps;
    if (!rc && h_adj0) {
        el_adj_unscale_kernel<<<d.ngrid, blk, 0, c->stream>>>(d_tab, 1);
        if (hipGetLastError() != hipSuccess) rc = B2R_E_HIP;
    }
    if (!rc) rc = el_grad_out(s, grd, adj, h_glam, h_gmu, h_gb, h_adj0);
    if (!rc) *misfit = 0.5 * ms_h;
    release();
    return rc;
}

// Batched reverse-sweep gradient of nshots shots (one block each, R = 1,
// frozen faces, sizes may differ).  Forward: batched steps, the state saved
// every ckpt steps.  Reverse: each step is undone (the scheme is
// time-reversible) together with its adjoint, the state reset from the
// checkpoint at every segment boundary; no per-step storage.  graph != 0
// replays whole segments from HIP graphs (step counters on the device).
// Device memory el_gradient_rev takes for one shot: checkpoints, adjoint
// state and gradients, observations and residuals, FP64 partial sums.
static size_t el_gradient_bytes(const b2r_el_shot* s, int nsteps, int ckpt)
{
    const int ncp = (nsteps + ckpt - 1) / ckpt;
    const size_t fe = (size_t)s->ctx->L.field_elems;
    const long long nobs = (long long)s->nrec * s->rec_nt;
    return ((size_t)9 * (ncp - 1) + 12) * fe * 4 + (size_t)nobs * 8 + 8 + 257 * sizeof(double);
}

static int el_gradient_rev(b2r_el_shot* const* shots, int nshots, const float* const* h_obs,
                           int nsteps, int ckpt, bool graph, int flags, double* misfit,
                           float* const* h_glam, float* const* h_gmu, float* const* h_gb,
                           float* h_adj0, float* ms)
{
    B2R_CHECK(shots && h_obs && misfit && h_glam && h_gmu && h_gb && nshots >= 1 &&
                  nshots <= EL_MAX_BATCH && nsteps >= 1 && ckpt >= 1,
              B2R_E_ARG, "bad argument");
    B2R_CHECK(!h_adj0 || nshots == 1, B2R_E_ARG, "the adjoint state is returned for one shot");
    for (int k = 0; k < nshots; ++k) {
        if (int rc = el_grad_check(shots[k], nsteps, true)) return rc;
        B2R_CHECK(h_obs[k] && h_glam[k] && h_gmu[k] && h_gb[k], B2R_E_ARG, "null array, shot %d", k);
    }
    b2r_ctx* c0 = shots[0]->ctx;
    if (int rc = b2r_activate(c0)) return rc;
    hipStream_t st = c0->stream;
    const int ncp = (nsteps + ckpt - 1) / ckpt;
    // per shot: [checkpoints (ncp-1) x 9][adjoint 9][gradient 3] fields, then
    // [obs][res] samples and FP64 partials
    std::vector<void*> mem(nshots, nullptr);
    std::vector<double> h_mis(nshots, 0.0);
    ElBatch bt;
    auto release = [&]() {
        el_batch_free(bt);
        (void)hipStreamSynchronize(st);
        for (void* p : mem)
            if (p) (void)hipFree(p);
        for (int k = 0; k < nshots; ++k) {
            shots[k]->d.cpt = nullptr;
            shots[k]->d.res = nullptr;
            for (int i = 0; i < 9; ++i) shots[k]->d.adj[i] = nullptr;
            for (int i = 0; i < 3; ++i) shots[k]->d.grad[i] = nullptr;
        }
    };
    int rc = B2R_OK;
    for (int k = 0; k < nshots && !rc; ++k) {
        b2r_el_shot* s = shots[k];
        const size_t fe = (size_t)s->ctx->L.field_elems;
        const long long nobs = (long long)s->nrec * s->rec_nt;
        const size_t bytes = el_gradient_bytes(s, nsteps, ckpt);
        if (hipMalloc(&mem[k], bytes) != hipSuccess) {
            mem[k] = nullptr;
            b2r_set_error("gradient: device allocation of %zu MB for shot %d failed", bytes >> 20, k);
            rc = B2R_E_NOMEM;
            break;
        }
        float* base = (float*)mem[k];
        s->d.cpt = base;
        s->d.fstride = (long long)fe;
        float* adj = base + (size_t)9 * (ncp - 1) * fe;
        for (int i = 0; i < 9; ++i) s->d.adj[i] = adj + i * fe;
        for (int i = 0; i < 3; ++i) s->d.grad[i] = adj + (9 + i) * fe;
        float* obs = adj + 12 * fe;
        s->d.res = obs + nobs;
        if (hipMemsetAsync(adj, 0, 12 * fe * 4, st) != hipSuccess) rc = B2R_E_HIP;
        if (!rc) {
            if (flags & 1) s->t = 0;
            else rc = b2r_el_reset(s);
        }
    }
    if (!rc) rc = el_batch_init(bt, shots, nshots, false);
    // the adjoint-propagation table (same workgroup numbering as bt.d_tab)
    ElDev* d_atab = nullptr;
    if (!rc) {
        std::vector<ElDev> at(nshots);
        int t0 = 0;
        for (int k = 0; k < nshots; ++k) {
            at[k] = el_adj_entry(shots[k]->d);
            at[k].tile0 = t0;
            t0 += shots[k]->d.ngrid;
        }
        rc = el_upload_table(at, st, &d_atab);
    }
    if (!rc) rc = b2r_time_begin(c0, ms);
    const dim3 blk(EL_TX, EL_TY);
    // ---- forward: segments of ckpt steps, each followed by a checkpoint save
    auto fwd_segment = [&](int n) {
        for (int i = 0; i < n; ++i)
            if (int r = el_group(bt, 1, false)) return r;
        el_ckpt_kernel<<<bt.total_tiles, blk, 0, st>>>(bt.d_tab, nshots, bt.d_t, ckpt, nsteps, 0);
        return hipGetLastError() == hipSuccess ? B2R_OK : B2R_E_HIP;
    };
    // ---- reverse: n steps, after a checkpoint restore at the segment top
    auto rev_segment = [&](int n) {
        el_ckpt_kernel<<<bt.total_tiles, blk, 0, st>>>(bt.d_tab, nshots, bt.d_t, ckpt, nsteps, 1);
        for (int i = 0; i < n; ++i) {
            if (bt.max_rec)
                el_adj_point_kernel<<<dim3(ceil_div(bt.max_rec, 64), nshots), 64, 0, st>>>(
                    bt.d_tab, bt.d_t, 0, 1);
            if (bt.max_src)
                el_adj_point_kernel<<<dim3(ceil_div(bt.max_src, 64), nshots), 64, 0, st>>>(
                    bt.d_tab, bt.d_t, 1, 1);
            // halos of the state after the step and of the adjoint, from their owners
            if (int r = el_exchange(bt)) return r;
            for (b2r_el_shot* s : bt.shots) {
                if (!el_grad_exchanges(s->ctx)) continue;
                hipStream_t keep = s->ctx->stream;
                s->ctx->stream = st;
                const int r = b2r_direct_exchange(s->ctx, (void* const*)s->d.adj, 9);
                s->ctx->stream = keep;
                if (r) return r;
            }
            el_adj_vel_kernel<true><<<bt.total_tiles, blk, 0, st>>>(bt.d_tab, nshots);
            el_stress_kernel<<<bt.total_tiles, blk, 0, st>>>(d_atab, nshots, 0);
            el_adj_str_kernel<true><<<bt.total_tiles, blk, 0, st>>>(bt.d_tab, nshots);
            el_velocity_kernel<7><<<bt.total_tiles, blk, 0, st>>>(d_atab, nshots, 1);
            el_tick_kernel<<<1, 64, 0, st>>>(bt.d_t, nshots, -1);
        }
        return hipGetLastError() == hipSuccess ? B2R_OK : B2R_E_HIP;
    };
    // a segment graph, captured once and replayed (its counters advance on the device)
    auto run_graph = [&](bool reverse, int count) {
        if (count <= 0) return B2R_OK;
        hipGraph_t g = nullptr;
        hipGraphExec_t ex = nullptr;
        int r = B2R_OK;
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            b2r_set_error("hipStreamBeginCapture failed");
            return B2R_E_HIP;
        }
        r = reverse ? rev_segment(ckpt) : fwd_segment(ckpt);
        hipError_t e = hipStreamEndCapture(st, &g);
        if (!r && e != hipSuccess) r = B2R_E_HIP;
        if (!r && hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) r = B2R_E_HIP;
        for (int i = 0; i < count && !r; ++i)
            if (hipGraphLaunch(ex, st) != hipSuccess) r = B2R_E_HIP;
        if (ex) (void)hipGraphExecDestroy(ex);
        if (g) (void)hipGraphDestroy(g);
        if (r) b2r_set_error("gradient: graph capture or launch failed");
        return r;
    };
    const int nfull = nsteps / ckpt, tail = nsteps - nfull * ckpt;
    if (!rc) {
        if (graph) rc = run_graph(false, nfull);
        else
            for (int j = 0; j < nfull && !rc; ++j) rc = fwd_segment(ckpt);
    }
    if (!rc && tail) rc = fwd_segment(tail);
    for (int k = 0; k < nshots && !rc; ++k) {
        b2r_el_shot* s = shots[k];
        const long long nobs = (long long)s->nrec * s->rec_nt;
        float* obs = (float*)s->d.res - nobs;
        double* part = (double*)(s->d.res + nobs + (nobs & 1));
        rc = el_residuals(s, st, h_obs[k], obs, (float*)s->d.res, part, &h_mis[k]);
    }
    // ---- reverse sweep: the top segment (the tail, or a full one) starts from
    // the final state, every other one from its checkpoint
    if (!rc) {
        el_tick_kernel<<<1, 64, 0, st>>>(bt.d_t, nshots, -1);
        rc = rev_segment(tail ? tail : ckpt);
    }
    const int nrest = tail ? nfull : nfull - 1;
    if (!rc) {
        if (graph) rc = run_graph(true, nrest);
        else
            for (int j = 0; j < nrest && !rc; ++j) rc = rev_segment(ckpt);
    }
    if (!rc) rc = b2r_time_end(c0, ms);
    if (!rc && h_adj0) {
        el_adj_unscale_kernel<<<bt.total_tiles, blk, 0, st>>>(bt.d_tab, nshots);
        if (hipGetLastError() != hipSuccess) rc = B2R_E_HIP;
    }
    if (d_atab) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(d_atab);
    }
    if (!rc) B2R_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < nshots && !rc; ++k) {
        b2r_el_shot* s = shots[k];
        rc = el_grad_out(s, s->d.grad[0], s->d.adj[0], h_glam[k], h_gmu[k], h_gb[k],
                         k == 0 ? h_adj0 : nullptr);
        misfit[k] = 0.5 * h_mis[k];
        s->t = nsteps;
    }
    release();
    return rc;
}

extern "C" int b2r_el_gradient_ex(b2r_el_shot* s, const float* h_obs, int nsteps, int ckpt,
                                  int flags, double* misfit, float* h_glam, float* h_gmu,
                                  float* h_gb, float* h_adj0)
{
    B2R_CHECK(s && h_obs && misfit && h_glam && h_gmu && h_gb && nsteps >= 1 && ckpt >= 1,
              B2R_E_ARG, "bad argument");
    if (flags & 2)
        return el_gradient_plain(s, h_obs, nsteps, ckpt, flags, misfit, h_glam, h_gmu, h_gb,
                                 h_adj0);
    return el_gradient_rev(&s, 1, &h_obs, nsteps, ckpt, false, flags, misfit, &h_glam, &h_gmu,
                           &h_gb, h_adj0, nullptr);
}

extern "C" int b2r_el_gradient(b2r_el_shot* s, const float* h_obs, int nsteps, int ckpt,
                               double* misfit, float* h_glam, float* h_gmu, float* h_gb)
{
    return b2r_el_gradient_ex(s, h_obs, nsteps, ckpt, 0, misfit, h_glam, h_gmu, h_gb, nullptr);
}

// The device-memory budget of the batched gradient (b2r_tuning.el_grad_mb of
// the first shot, 0 = the free memory less 5 %).
static int el_grad_budget(b2r_el_shot* const* shots, size_t* budget)
{
    if (int rc = b2r_activate(shots[0]->ctx)) return rc;
    *budget = (size_t)shots[0]->ctx->tune.el_grad_mb << 20;
    if (*budget == 0) {
        size_t free_b = 0, total_b = 0;
        B2R_HIP(hipMemGetInfo(&free_b, &total_b));
        *budget = free_b / 20 * 19;
    }
    return B2R_OK;
}

extern "C" int b2r_el_gradient_plan(b2r_el_shot* const* shots, int nshots, int nsteps, int ckpt,
                                    int* ngroups, size_t* bytes_max_group)
{
    B2R_CHECK(shots && nshots >= 1 && nsteps >= 1 && ckpt >= 1 && ngroups, B2R_E_ARG,
              "bad argument");
    for (int k = 0; k < nshots; ++k) B2R_CHECK(shots[k], B2R_E_ARG, "null shot %d", k);
    size_t budget = 0;
    if (int rc = el_grad_budget(shots, &budget)) return rc;
    int groups = 0;
    size_t big = 0;
    for (int k0 = 0; k0 < nshots; ++groups) {
        size_t used = 0;
        int k1 = k0;
        while (k1 < nshots) {
            const size_t need = el_gradient_bytes(shots[k1], nsteps, ckpt);
            if (k1 > k0 && used + need > budget) break;
            B2R_CHECK(need <= budget, B2R_E_NOMEM,
                      "gradient: shot %d needs %zu MB, more than the %zu MB allowed", k1,
                      need >> 20, budget >> 20);
            used += need;
            ++k1;
        }
        big = used > big ? used : big;
        k0 = k1;
    }
    *ngroups = groups;
    if (bytes_max_group) *bytes_max_group = big;
    return B2R_OK;
}

extern "C" int b2r_el_gradient_batch(b2r_el_shot* const* shots, int nshots,
                                     const float* const* h_obs, int nsteps, int ckpt, int mode,
                                     double* misfit, float* const* h_glam, float* const* h_gmu,
                                     float* const* h_gb, float* ms)
{
    B2R_CHECK(shots && h_obs && misfit && h_glam && h_gmu && h_gb && nshots >= 1 &&
                  mode >= 0 && mode <= 2 && nsteps >= 1 && ckpt >= 1,
              B2R_E_ARG, "bad argument");
    if (mode > 0) {
        // consecutive groups of shots whose gradient memory fits the budget
        for (int k = 0; k < nshots; ++k) B2R_CHECK(shots[k], B2R_E_ARG, "null shot %d", k);
        size_t budget = 0;
        if (int rc = el_grad_budget(shots, &budget)) return rc;
        float total_ms = 0.0f;
        for (int k0 = 0; k0 < nshots;) {
            int k1 = k0;
            size_t used = 0;
            while (k1 < nshots) {
                const size_t need = el_gradient_bytes(shots[k1], nsteps, ckpt);
                if (k1 > k0 && used + need > budget) break;
                B2R_CHECK(need <= budget, B2R_E_NOMEM,
                          "gradient: shot %d needs %zu MB, more than the %zu MB allowed", k1,
                          need >> 20, budget >> 20);
                used += need;
                ++k1;
            }
            float gms = 0.0f;
            if (int rc = el_gradient_rev(shots + k0, k1 - k0, h_obs + k0, nsteps, ckpt, mode == 2, 0,
                                         misfit + k0, h_glam + k0, h_gmu + k0, h_gb + k0, nullptr,
                                         ms ? &gms : nullptr))
                return rc;
            total_ms += gms;
            k0 = k1;
        }
        if (ms) *ms = total_ms;
        return B2R_OK;
    }
    // plain: one shot after the other, timed as a whole
    if (int rc = b2r_activate(shots[0]->ctx)) return rc;
    if (int rc = b2r_time_begin(shots[0]->ctx, ms)) return rc;
    for (int k = 0; k < nshots; ++k) {
        B2R_CHECK(shots[k] && shots[k]->ctx->device == shots[0]->ctx->device, B2R_E_ARG,
                  "all shots of a batch must be on one device");
        if (int rc = el_gradient_plain(shots[k], h_obs[k], nsteps, ckpt, 0, &misfit[k],
                                       h_glam[k], h_gmu[k], h_gb[k], nullptr))
            return rc;
    }
    return b2r_time_end(shots[0]->ctx, ms);
}
