// b2r_heat.hip -- heat diffusion fdm_heat_diffuse_delta1..4 for gfx950.
//
// Reference: heat-3D.cu:173-521 (kernels), :583-642 (R-step driver),
// heat-3D.cpp:48-62 (CPU step).  The reference kernels use 32x32 thread blocks,
// load tile-edge halos with edge threads and march z with a register queue; they
// read the y-halo from the wrong row for D>1 and start the z march at the wrong
// plane for R>1, which is not reproduced.
//
// This kernel: one workgroup = 64*WX x-lanes x WY thread rows; each thread
// owns RY consecutive rows of one x column and marches z through a chunk of
// planes.  The z neighbours live in registers: for D <= 2 a queue of 2D+1
// values per row shifted by one plane per step, for D >= 3 a ring of 2D+2
// planes (the march unrolled by the ring length, so no value moves, and the
// loads are buffer loads whose plane offset is a scalar).  The current plane,
// with a D-wide x/y halo, is staged in LDS so the x and y neighbours are LDS
// reads (two LDS planes alternate, one barrier per plane).  Loads of the
// z-front are coalesced rows, issued one plane ahead.
//
// Arithmetic (compiled without FP contraction, so CPU and GPU round alike):
//   F32: u' = u + ((kx*Lx + ky*Ly) + kz*Lz)
//   REF: u' = (float)((double)u + adt*(((double)Lx/hx2 + (double)Ly/hy2) + (double)Lz/hz2))
//   L_a = c_D*(m_D+p_D) + ... + c_1*(m_1+p_1) + c_0*u   (heat-3D.cu:314-318 order),
//   REF and D=1: L_a = (m_1 - 2u) + p_1                  (heat-3D.cpp:58-60).
#include "b2r_common.h"
#include "b2r_heat_ops.h"

#include <type_traits>

constexpr int HEAT_TX = 64;

constexpr int HEAT_WY = 4;  // thread rows (waves) of the fused kernel

template <int D, int MODE, int RY, int WY, int WX>
__global__ __launch_bounds__(HEAT_TX * WX * WY) void heat_step_kernel(
    const float* __restrict__ in, float* __restrict__ out, KGeo g, HeatK hk, int b2r_i, int ntx,
    int nty, int nzc, int zc)
{
    constexpr int TX = HEAT_TX * WX, TY = WY * RY, NT = TX * WY;
    constexpr int LX = TX + 2 * D, LY = TY + 2 * D;
    constexpr int NHY = 2 * D * TX, NHX = 2 * D * TY, NH = NHY + NHX;
    constexpr int NSLOT = (NH + NT - 1) / NT;
    // two staged planes, alternating: the plane written at z was last read at
    // z-2, before every thread passed the barrier of z-1, so one barrier per
    // plane separates the writes from the reads
    __shared__ float lds2[2][LY * LX];

    const int lane = threadIdx.x, wy = threadIdx.y, tid = wy * TX + lane;
    int t = xcd_tile(blockIdx.x, gridDim.x);
    const int txi = t % ntx;
    t /= ntx;
    const int tyi = t % nty;
    t /= nty;
    const int zci = t % nzc;
    const int b = t / nzc;

    const Box box = update_box(g, b, D, b2r_i);
    const int x0 = D + txi * TX, y0 = D + tyi * TY;
    if (x0 > box.hi[0] || x0 + TX - 1 < box.lo[0] || y0 > box.hi[1] || y0 + TY - 1 < box.lo[1])
        return;  // uniform over the workgroup
    const int zs = D + zci * zc;
    const int ze = min(zs + zc, g.S[2] - D) - 1;
    const int zlo = max(zs, box.lo[2]), zhi = min(ze, box.hi[2]);
    if (zlo > zhi) return;

    // the workgroup reads planes zlo-D .. zhi+D: one base pointer at plane
    // zlo-D, cells addressed by 32-bit byte offsets from it (the launch keeps
    // zc + 2D planes below 4 GiB)
    const long long base = (long long)b * g.bstride + g.xoff + (long long)(zlo - D) * g.slice;
    const uint32_t sz = 4u * (uint32_t)g.slice;
    const uint32_t pitch = 4u * (uint32_t)g.pitch;

    const int x = x0 + lane;
    const bool xin = x < g.S[0];
    const bool xupd = x >= box.lo[0] && x <= box.hi[0];
    int y[RY];
    // byte offset of the thread's cell of row j within a plane; 0 (the
    // plane's first cell) for a cell outside the block, so that the loads of
    // the march are unconditional: the value read there is never used, and
    // stores happen only inside the update box, which lies in the block
    uint32_t yo[RY];
    bool yin[RY], yupd[RY];
#pragma unroll
    for (int j = 0; j < RY; ++j) {
        y[j] = y0 + wy * RY + j;
        yin[j] = y[j] < g.S[1];
        yupd[j] = y[j] >= box.lo[1] && y[j] <= box.hi[1];
        yo[j] = (xin && yin[j]) ? (uint32_t)y[j] * pitch + 4u * (uint32_t)x : 0u;
    }

    // halo slots: D rows above/below the tile (TX wide), D columns left/right (TY tall)
    int hl[NSLOT];
    uint32_t ho[NSLOT];
    bool hok[NSLOT];
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
        const int e = tid + s * NT;
        hok[s] = false;
        hl[s] = 0;
        ho[s] = 0;
        if (e < NH) {
            int lx, ly;
            if (e < NHY) {
                const int r = e / TX, c = e % TX;
                ly = r < D ? r : TY + r;
                lx = c + D;
            } else {
                const int e2 = e - NHY;
                const int r = e2 % TY, c = e2 / TY;
                ly = r + D;
                lx = c < D ? c : TX + c;
            }
            const int gx = x0 - D + lx, gy = y0 - D + ly;
            hok[s] = gx < g.S[0] && gy < g.S[1];
            hl[s] = ly * LX + lx;
            ho[s] = hok[s] ? (uint32_t)gy * pitch + 4u * (uint32_t)gx : 0u;
        }
    }

    if constexpr (D <= 2) {
        // D = 1, 2: the z-queue of 2D+1 values per row shifts by one plane
        // per step.  (The ring below issues a third fewer VALU instructions,
        // which pays where the stencil's arithmetic is large; at D = 1 and 2
        // the shifted queue measured 7 % and 1.5 % faster.)
        const gfloat* __restrict__ src = G(in + base);
        gfloat* __restrict__ dst = G(out + base);
        constexpr int Q = 2 * D + 1;
        float q[RY][Q];
    #pragma unroll
        for (int p = 0; p < 2 * D; ++p) {
    #pragma unroll
            for (int j = 0; j < RY; ++j)
                q[j][p] = (xin && yin[j]) ? ldo(src, p * sz + yo[j]) : 0.0f;
        }

        // The tile's values are loaded D planes ahead of the next plane (qn, for
        // the z-queue).  For D >= 2 so are the halo values (a queue hq of HA+1
        // planes): a halo row is its neighbour tile's own row, which that tile
        // loads D planes ahead, so both requests reach L2 at about the same time
        // instead of D planes apart (about 1 MB of plane data per XCD per plane at
        // D=4, which the 4 MB L2 does not keep for 4 planes).  At D=1 the halo is
        // loaded one plane ahead.
        constexpr int HA = D >= 2 ? D : 0;
        float hq[NSLOT][HA + 1], qn[RY];
    #pragma unroll
        for (int s = 0; s < NSLOT; ++s)
    #pragma unroll
            for (int k = 0; k <= HA; ++k) hq[s][k] = hok[s] ? ldo(src, (D + k) * sz + ho[s]) : 0.0f;
    #pragma unroll
        for (int j = 0; j < RY; ++j) qn[j] = (xin && yin[j]) ? ldo(src, 2 * D * sz + yo[j]) : 0.0f;
        // one plane of the march; lds is lds2[0] or lds2[1] as a constant of each
        // call, so the LDS addresses stay loop-invariant
        auto plane = [&](const int z, const uint32_t pz, float* const lds) {
    #pragma unroll
            for (int j = 0; j < RY; ++j) q[j][2 * D] = qn[j];
    #pragma unroll
            for (int j = 0; j < RY; ++j) lds[(wy * RY + j + D) * LX + lane + D] = q[j][D];
    #pragma unroll
            for (int s = 0; s < NSLOT; ++s) {
                if (hok[s]) lds[hl[s]] = hq[s][0];
    #pragma unroll
                for (int k = 0; k < HA; ++k) hq[s][k] = hq[s][k + 1];
            }
            if (z < zhi) {
                // planes z+1+HA and z+1+D: z+1 <= zhi <= S[2]-1-D keeps both in the block
                const uint32_t pn = pz + (D + 1) * sz;
    #pragma unroll
                for (int j = 0; j < RY; ++j)
                    if (xin && yin[j]) qn[j] = ldo(src, pn + yo[j]);
    #pragma unroll
                for (int s = 0; s < NSLOT; ++s)
                    if (hok[s]) hq[s][HA] = ldo(src, pz + (HA + 1) * sz + ho[s]);
            }
            __syncthreads();
    #pragma unroll
            for (int j = 0; j < RY; ++j) {
                if (xupd && yupd[j]) {
                    const float* row = &lds[(wy * RY + j + D) * LX + lane + D];
                    const float uc = q[j][D];
                    float mx[D + 1], px[D + 1], my[D + 1], py[D + 1], mz[D + 1], pz_[D + 1];
    #pragma unroll
                    for (int k = 1; k <= D; ++k) {
                        mx[k] = row[-k];
                        px[k] = row[k];
                        // rows of this thread are in registers (j, k are unrolled)
                        my[k] = j - k >= 0 ? q[j - k >= 0 ? j - k : 0][D] : row[-k * LX];
                        py[k] = j + k < RY ? q[j + k < RY ? j + k : 0][D] : row[k * LX];
                        mz[k] = q[j][D - k];
                        pz_[k] = q[j][D + k];
                    }
                    sto(dst, pz + yo[j], heat_update<D, MODE>(mx, px, my, py, mz, pz_, uc, hk));
                }
            }
    #pragma unroll
            for (int j = 0; j < RY; ++j)
    #pragma unroll
                for (int p = 0; p < 2 * D; ++p) q[j][p] = q[j][p + 1];
        };
        // planes in pairs (the buffer of plane z is (z - zlo) & 1), an odd one last
        int z = zlo;
        uint32_t pz = D * sz;  // byte offset of plane z
        for (; z + 1 <= zhi; z += 2, pz += 2 * sz) {
            plane(z, pz, lds2[0]);
            plane(z + 1, pz + sz, lds2[1]);
        }
        if (z <= zhi) plane(z, pz, lds2[0]);
    } else {
        const __amdgpu_buffer_rsrc_t src = buf_rsrc(in + base);
        const __amdgpu_buffer_rsrc_t dst = buf_rsrc(out + base);
        // The z-queue is a ring of L = 2D+2 planes per row: at ring step r
        // (plane z of the march, r = (z - zlo) mod L) slot (r + D + d) mod L holds
        // plane z + d for d = -D..D, and slot (r + 2D + 1) mod L, the one plane
        // z-D-1 left, receives plane z+D+1, loaded one plane ahead of its use.
        // The march is unrolled by L, so every slot index is a compile-time
        // constant: no value moves between registers (a shifted queue costs 2D
        // moves per cell) and a load in flight is never copied, so it is waited
        // for only where it is used.  L is even, so the LDS plane of step r is
        // lds2[r & 1] across iterations as well.
        //   The halo values of plane z are loaded HA+1 planes ahead into a ring of
        // L slots per halo slot (plane z + d in slot (r + d) mod L).  For D >= 2,
        // HA = D: a halo row is its neighbour tile's own row, which that tile
        // loads D+1 planes ahead, so both requests reach L2 at about the same time
        // instead of D planes apart (about 1 MB of plane data per XCD per plane at
        // D=4, which the 4 MB L2 does not keep for 4 planes).  At D=1, HA = 0.
        constexpr int L = 2 * D + 2;
        constexpr int HA = D >= 2 ? D : 0;
        static_assert(HA + 1 < L, "halo ring too short");
        float q[RY][L], hq[NSLOT][L];
        // planes zlo-D .. zlo+D into slots 0 .. 2D (ring step 0).  A cell
        // outside the block reads the plane's first cell instead: no updated cell
        // reads it (an updated cell's stencil lies in the block), so the loads
        // are unconditional, as in the march.
    #pragma unroll
        for (int p = 0; p <= 2 * D; ++p) {
    #pragma unroll
            for (int j = 0; j < RY; ++j) q[j][p] = bld(src, yo[j], p * sz);
        }
    #pragma unroll
        for (int s = 0; s < NSLOT; ++s)
    #pragma unroll
            for (int k = 0; k <= HA; ++k) hq[s][k] = bld(src, ho[s], (D + k) * sz);
        // wait for all of them here, before the march: the waits at the loop's
        // head serve the loop's entry as well, and without this one they would
        // wait for the last prologue loads on every iteration (vmcnt(0),
        // expcnt and lgkmcnt left open: 0x0F70 on gfx9)
        __builtin_amdgcn_s_waitcnt(0x0F70);

        // ring step r at plane z, whose byte offset from the base plane is pz
        auto step = [&](auto rc, const int z, const uint32_t pz) {
            constexpr int r = decltype(rc)::value;
            constexpr int C = (r + D) % L;  // slot of plane z
            float* const lds = lds2[r & 1];
    #pragma unroll
            for (int j = 0; j < RY; ++j) lds[(wy * RY + j + D) * LX + lane + D] = q[j][C];
    #pragma unroll
            for (int s = 0; s < NSLOT; ++s)
                if (hok[s]) lds[hl[s]] = hq[s][r % L];
            {
                // halo plane z+1+HA and tile plane z+1+D.  The last plane, zhi,
                // loads planes zhi+HA and zhi+D again instead (values unused),
                // keeping the loads unconditional and inside the block (zhi+D <=
                // S[2]-1).  The plane offset is the scalar offset of the buffer
                // access, so advancing a plane adds nothing to the per-thread
                // offsets.
                const uint32_t adv = z < zhi ? sz : 0u;
    #pragma unroll
                for (int s = 0; s < NSLOT; ++s) hq[s][(r + HA + 1) % L] = bld(src, ho[s], pz + HA * sz + adv);
    #pragma unroll
                for (int j = 0; j < RY; ++j) q[j][(r + 2 * D + 1) % L] = bld(src, yo[j], pz + D * sz + adv);
            }
            __syncthreads();
    #pragma unroll
            for (int j = 0; j < RY; ++j) {
                if (xupd && yupd[j]) {
                    const float* row = &lds[(wy * RY + j + D) * LX + lane + D];
                    const float uc = q[j][C];
                    float mx[D + 1], px[D + 1], my[D + 1], py[D + 1], mz[D + 1], pz_[D + 1];
    #pragma unroll
                    for (int k = 1; k <= D; ++k) {
                        mx[k] = row[-k];
                        px[k] = row[k];
                        // rows of this thread are in registers (j, k are unrolled)
                        my[k] = j - k >= 0 ? q[j - k >= 0 ? j - k : 0][C] : row[-k * LX];
                        py[k] = j + k < RY ? q[j + k < RY ? j + k : 0][C] : row[k * LX];
                        mz[k] = q[j][(r + D - k + L) % L];
                        pz_[k] = q[j][(r + D + k) % L];
                    }
                    bst(dst, yo[j], pz, heat_update<D, MODE>(mx, px, my, py, mz, pz_, uc, hk));
                }
            }
        };
        // L planes per iteration; the kernel ends after plane zhi.  Each step
        // leaves the kernel itself, so no path that skips the later steps of an
        // iteration reaches the loop's back edge: the wait analysis, which merges
        // paths, then sees every step's loads issued before the next iteration.
        int z = zlo;
        uint32_t pz = D * sz;  // byte offset of plane z
    #define B2R_HEAT_STEP(R)                                        \
        step(std::integral_constant<int, R>{}, z, pz);              \
        if (++z > zhi) return;                                      \
        pz += sz;
        for (;;) {
            B2R_HEAT_STEP(0)
            B2R_HEAT_STEP(1)
            B2R_HEAT_STEP(2)
            B2R_HEAT_STEP(3)
            if constexpr (L > 4) {
                B2R_HEAT_STEP(4)
                B2R_HEAT_STEP(5)
            }
            if constexpr (L > 6) {
                B2R_HEAT_STEP(6)
                B2R_HEAT_STEP(7)
            }
            if constexpr (L > 8) {
                B2R_HEAT_STEP(8)
                B2R_HEAT_STEP(9)
            }
        }
    #undef B2R_HEAT_STEP
        static_assert(L == 4 || L == 6 || L == 8 || L == 10, "ring of 2D+2 planes, D = 1..4");
    }
}

int b2r_heat_check(b2r_ctx* ctx, const b2r_heat_params* p)
{
    B2R_CHECK(ctx && p, B2R_E_ARG, "null argument");
    B2R_CHECK(ctx->g.elem_size == 4, B2R_E_UNSUPPORTED, "heat needs f32 cells");
    B2R_CHECK(p->D >= 1 && p->D <= 4, B2R_E_UNSUPPORTED, "unsupported delta D=%d", p->D);
    B2R_CHECK(p->D == ctx->g.radius, B2R_E_ARG, "heat D=%d but context radius=%d", p->D,
              ctx->g.radius);
    B2R_CHECK(p->mode == B2R_HEAT_F32 || p->mode == B2R_HEAT_REF, B2R_E_ARG, "bad mode");
    // the step kernel addresses 2D + 1 planes or more by 32-bit byte offsets
    B2R_CHECK(4ull * (unsigned long long)ctx->k.slice * (2 * p->D + 2) <= 0xffffffffull,
              B2R_E_UNSUPPORTED, "x-y plane too large for the heat kernels");
    if (p->mode == B2R_HEAT_REF)
        B2R_CHECK(p->h2[0] > 0 && p->h2[1] > 0 && p->h2[2] > 0, B2R_E_ARG, "h2 must be > 0");
    // a 2D grid has no z term: its z frame is zero, so the step must not see
    // a z difference (F32 mode with k_z = 0; the reference expression of
    // heat-3D.cpp:57-60 always has one)
    if (ctx->k.two_d)
        B2R_CHECK(p->mode == B2R_HEAT_F32 && p->k[2] == 0.0f, B2R_E_UNSUPPORTED,
                  "a 2D grid (env_z = 1) runs F32 mode with k[2] = 0");
    return B2R_OK;
}

template <int D, int MODE, int RY, int WY, int WX>
static void launch_heat_ry(b2r_ctx* ctx, const float* in, float* out, const HeatK& hk, int i)
{
    constexpr int TX = HEAT_TX * WX, TY = WY * RY;
    const KGeo& g = ctx->k;
    const int ntx = ceil_div(g.S[0] - 2 * D, TX);
    const int nty = ceil_div(g.S[1] - 2 * D, TY);
    const int nz = g.S[2] - 2 * D;
    int zc = ctx->tune.heat_zc;
    if (zc <= 0) {
        // enough workgroups to fill 256 CUs several times, chunks >= 4D planes
        long long tiles = (long long)ntx * nty * g.nblk;
        int nzc = (int)((4096 + tiles - 1) / tiles);
        zc = ceil_div(nz, nzc < 1 ? 1 : nzc);
        if (zc < 4 * D) zc = 4 * D;
    }
    if (zc > nz) zc = nz;
    // a workgroup addresses its zc + 2D planes by 32-bit byte offsets
    const long long zmax = (long long)(0xffffffffu / (4ull * (unsigned long long)g.slice)) - 2 * D;
    if (zc > zmax) zc = (int)zmax;
    const int nzc = ceil_div(nz, zc);
    dim3 grid(ntx * nty * nzc * g.nblk);
    dim3 block(TX, WY);
    heat_step_kernel<D, MODE, RY, WY, WX><<<grid, block, 0, ctx->stream>>>(in, out, g, hk, i, ntx,
                                                                           nty, nzc, zc);
}

template <int D, int MODE, int WY, int WX>
static void launch_heat_wy(b2r_ctx* ctx, const float* in, float* out, const HeatK& hk, int i)
{
    switch (ctx->tune.heat_ry) {
    case 1: launch_heat_ry<D, MODE, 1, WY, WX>(ctx, in, out, hk, i); break;
    case 2: launch_heat_ry<D, MODE, 2, WY, WX>(ctx, in, out, hk, i); break;
    case 8: launch_heat_ry<D, MODE, 8, WY, WX>(ctx, in, out, hk, i); break;
    default: launch_heat_ry<D, MODE, 4, WY, WX>(ctx, in, out, hk, i); break;
    }
}

// workgroup = 64*heat_wx x-lanes x heat_wy thread rows, each thread heat_ry
// rows.  (heat_wy, heat_wx): (4, 1) by default, (4, 2), (4, 4), (8, 1), (8, 2),
// (16, 1), the 12-wave shapes (4, 3) and (2, 6), whose 192- and 384-cell rows
// cover a 384-cell block row in 2 or 1 tiles, and the 9-wave (3, 3)
template <int D, int MODE>
static void launch_heat(b2r_ctx* ctx, const float* in, float* out, const HeatK& hk, int i)
{
    const int wx = ctx->tune.heat_wx;
    switch (ctx->tune.heat_wy) {
    case 2: launch_heat_wy<D, MODE, 2, 6>(ctx, in, out, hk, i); break;
    case 3: launch_heat_wy<D, MODE, 3, 3>(ctx, in, out, hk, i); break;
    case 8:
        if (wx == 2) launch_heat_wy<D, MODE, 8, 2>(ctx, in, out, hk, i);
        else launch_heat_wy<D, MODE, 8, 1>(ctx, in, out, hk, i);
        break;
    case 16: launch_heat_wy<D, MODE, 16, 1>(ctx, in, out, hk, i); break;
    default:
        if (wx == 4) launch_heat_wy<D, MODE, 4, 4>(ctx, in, out, hk, i);
        else if (wx == 3) launch_heat_wy<D, MODE, 4, 3>(ctx, in, out, hk, i);
        else if (wx == 2) launch_heat_wy<D, MODE, 4, 2>(ctx, in, out, hk, i);
        else launch_heat_wy<D, MODE, 4, 1>(ctx, in, out, hk, i);
        break;
    }
}

int b2r_heat_launch(b2r_ctx* ctx, const void* in, void* out, const b2r_heat_params* p, int i)
{
    const HeatK hk = heat_k(p);
    const float* fi = (const float*)in;
    float* fo = (float*)out;
#define B2R_HEAT_CASE(DD)                                                              \
    case DD:                                                                           \
        if (p->mode == B2R_HEAT_F32) launch_heat<DD, B2R_HEAT_F32>(ctx, fi, fo, hk, i); \
        else launch_heat<DD, B2R_HEAT_REF>(ctx, fi, fo, hk, i);                         \
        break;
    switch (p->D) {
        B2R_HEAT_CASE(1)
        B2R_HEAT_CASE(2)
        B2R_HEAT_CASE(3)
        B2R_HEAT_CASE(4)
    }
#undef B2R_HEAT_CASE
    B2R_LAUNCH_CHECK();
    return B2R_OK;
}

extern "C" int b2r_heat_step(b2r_ctx* ctx, const void* d_in, void* d_out,
                             const b2r_heat_params* p, int b2r_i)
{
    if (int rc = b2r_heat_check(ctx, p)) return rc;
    if (int rc = b2r_check_field(ctx, d_in)) return rc;
    if (int rc = b2r_check_field(ctx, d_out)) return rc;
    B2R_CHECK(d_in != d_out, B2R_E_ARG, "in and out must differ");
    B2R_CHECK(b2r_i >= 0 && b2r_i < ctx->g.R, B2R_E_ARG, "b2r_i=%d outside [0,R)", b2r_i);
    if (int rc = b2r_activate(ctx)) return rc;
    return b2r_heat_launch(ctx, d_in, d_out, p, b2r_i);
}

static bool heat_fused_ok(const b2r_ctx* ctx, int D);

extern "C" int b2r_heat_run(b2r_ctx* ctx, void* d_a, void* d_b, const b2r_heat_params* p,
                            int nsteps, int fused, int* result_in_b, float* ms)
{
    if (int rc = b2r_heat_check(ctx, p)) return rc;
    if (int rc = b2r_check_field(ctx, d_a)) return rc;
    if (int rc = b2r_check_field(ctx, d_b)) return rc;
    B2R_CHECK(d_a != d_b && nsteps >= 0, B2R_E_ARG, "bad buffers or nsteps");
    if (int rc = b2r_activate(ctx)) return rc;
    // both ping-pong buffers start with the same halos (heat-3D.cu:590-592)
    if (int rc = b2r_field_copy(ctx, d_b, d_a)) return rc;
    void* cur = d_a;
    void* nxt = d_b;
    if (int rc = b2r_time_begin(ctx, ms)) return rc;
    const int R = ctx->g.R;
    const bool use_fused = fused && heat_fused_ok(ctx, p->D);
    for (int s = 0; s < nsteps; s += R) {
        if (int rc = b2r_direct_exchange(ctx, &cur, 1)) return rc;
        const int n = nsteps - s < R ? nsteps - s : R;
        if (use_fused && n == R) {
            if (int rc = b2r_heat_fused_launch(ctx, cur, nxt, p)) return rc;
            void* t = cur;
            cur = nxt;
            nxt = t;
            continue;
        }
        for (int i = 0; i < n; ++i) {
            if (int rc = b2r_heat_launch(ctx, cur, nxt, p, i)) return rc;
            void* t = cur;
            cur = nxt;
            nxt = t;
        }
    }
    if (int rc = b2r_time_end(ctx, ms)) return rc;
    if (result_in_b) *result_in_b = cur == d_b;
    return B2R_OK;
}

// ---------------------------------------------------------------------------
// R steps in one launch: temporal blocking (B2R's deep halo kept on chip).
// A workgroup of 64 x 4 threads, RY = 8 rows each, covers a 64 x 32 column
// tile and marches z.  Level t (step t of the window, t = 1..R) of plane z_in-tD
// is computed from level t-1, whose last 2D+1 planes of the thread's own
// columns sit in a register queue; the x/y neighbours of the plane being
// computed come from one shared LDS plane (with a D-wide unused border),
// reused level after level.  The valid region shrinks by D per level, so the
// inner (64-2RD) x (32-2RD) columns of level R are written (R*D <= 8: at least
// 48 x 16 of the 64 x 32).  Cells outside a
// level's update box (frozen global faces) keep their previous value; halo
// cells of the output are never written.  Same arithmetic as the step kernel,
// so the results are bit-identical.
// ---------------------------------------------------------------------------
constexpr int HF_RY = 8;
constexpr int HF_TYT = HEAT_WY * HF_RY;  // thread-tile rows

template <int D, int MODE, int R>
__global__ __launch_bounds__(256) void heat_fused_kernel(const float* __restrict__ in,
                                                         float* __restrict__ out, KGeo g,
                                                         HeatK hk, int ntx, int nty, int nzc,
                                                         int zc)
{
    constexpr int RD = R * D;
    constexpr int TXO = HEAT_TX - 2 * RD, TYO = HF_TYT - 2 * RD;
    constexpr int LX = HEAT_TX + 2 * D, LY = HF_TYT + 2 * D;
    constexpr int Q = 2 * D + 1;
    __shared__ float lds[LY * LX];

    const int lane = threadIdx.x, wy = threadIdx.y;
    int tt = xcd_tile(blockIdx.x, gridDim.x);
    const int txi = tt % ntx;
    tt /= ntx;
    const int tyi = tt % nty;
    tt /= nty;
    const int zci = tt % nzc;
    const int b = tt / nzc;
    const int h = g.h;
    const int xo = h + txi * TXO;  // first output column
    const int yo = h + tyi * TYO;
    const int zc0 = h + zci * zc;
    const int zc1 = min(zc0 + zc, h + g.B[2]);

    const long long boff = (long long)b * g.bstride + g.xoff;
    const float* __restrict__ src = in + boff;
    float* __restrict__ dst = out + boff;
    const long long slice = g.slice;
    const int pitch = g.pitch;

    const int x = xo - RD + lane;
    const bool xin = x < g.S[0];
    int y[HF_RY];
    bool yin[HF_RY];
#pragma unroll
    for (int j = 0; j < HF_RY; ++j) {
        y[j] = yo - RD + wy * HF_RY + j;
        yin[j] = y[j] < g.S[1];
    }
    Box bx[R];
#pragma unroll
    for (int t = 0; t < R; ++t) bx[t] = update_box(g, b, D, t);

    // q[t][j][k]: level t at plane (newest - 2D + k) for row j
    float q[R][HF_RY][Q];
#pragma unroll
    for (int t = 0; t < R; ++t)
#pragma unroll
        for (int j = 0; j < HF_RY; ++j)
#pragma unroll
            for (int k = 0; k < Q; ++k) q[t][j][k] = 0.0f;

    const int lrow0 = wy * HF_RY + D;  // LDS row of this thread's first row
    for (int zin = zc0 - RD; zin < zc1 + RD; ++zin) {
        // level 0: push plane zin
        const long long pz = (long long)zin * slice;
#pragma unroll
        for (int j = 0; j < HF_RY; ++j) {
#pragma unroll
            for (int k = 0; k < Q - 1; ++k) q[0][j][k] = q[0][j][k + 1];
            q[0][j][Q - 1] = (xin && yin[j]) ? src[pz + (long long)y[j] * pitch + x] : 0.0f;
        }
#pragma unroll
        for (int t = 1; t <= R; ++t) {
            const int p = zin - t * D;
            const bool active = zin >= zc0 - RD + 2 * t * D;  // uniform
            __syncthreads();
#pragma unroll
            for (int j = 0; j < HF_RY; ++j) lds[(lrow0 + j) * LX + lane + D] = q[t - 1][j][D];
            __syncthreads();
            const Box& B_ = bx[t - 1];
            const bool zin_box = p >= B_.lo[2] && p <= B_.hi[2];
            const bool xin_box = x >= B_.lo[0] && x <= B_.hi[0];
#pragma unroll
            for (int j = 0; j < HF_RY; ++j) {
                const float uc = q[t - 1][j][D];
                float r = uc;
                if (active && zin_box && xin_box && y[j] >= B_.lo[1] && y[j] <= B_.hi[1]) {
                    const float* row = &lds[(lrow0 + j) * LX + lane + D];
                    float mx[D + 1], px[D + 1], my[D + 1], py[D + 1], mz[D + 1], pz_[D + 1];
#pragma unroll
                    for (int k = 1; k <= D; ++k) {
                        mx[k] = row[-k];
                        px[k] = row[k];
                        my[k] = row[-k * LX];
                        py[k] = row[k * LX];
                        mz[k] = q[t - 1][j][D - k];
                        pz_[k] = q[t - 1][j][D + k];
                    }
                    r = heat_update<D, MODE>(mx, px, my, py, mz, pz_, uc, hk);
                }
                if (t < R) {
#pragma unroll
                    for (int k = 0; k < Q - 1; ++k) q[t][j][k] = q[t][j][k + 1];
                    q[t][j][Q - 1] = r;
                } else if (active && lane >= RD && lane < HEAT_TX - RD && x < h + g.B[0]) {
                    const int lr = wy * HF_RY + j;
                    if (lr >= RD && lr < HF_TYT - RD && y[j] < h + g.B[1] && p < zc1)
                        dst[(long long)p * slice + (long long)y[j] * pitch + x] = r;
                }
            }
        }
    }
}

static bool heat_fused_ok(const b2r_ctx* ctx, int D)
{
    const int R = ctx->g.R;
    return R * D <= 8 && closed_faces(ctx->k);
}

template <int D, int MODE, int R>
static int heat_fused_t(b2r_ctx* ctx, const float* in, float* out, const HeatK& hk)
{
    constexpr int RD = R * D;
    constexpr int TXO = HEAT_TX - 2 * RD, TYO = HF_TYT - 2 * RD;
    const KGeo& g = ctx->k;
    const int ntx = ceil_div(g.B[0], TXO);
    const int nty = ceil_div(g.B[1], TYO);
    int zc = ctx->tune.heat_zc;
    if (zc <= 0) {
        long long tiles = (long long)ntx * nty * g.nblk;
        int nzc = (int)((2048 + tiles - 1) / tiles);
        zc = ceil_div(g.B[2], nzc < 1 ? 1 : nzc);
        if (zc < 8 * RD) zc = 8 * RD;
    }
    if (zc > g.B[2]) zc = g.B[2];
    const int nzc = ceil_div(g.B[2], zc);
    dim3 grid(ntx * nty * nzc * g.nblk);
    dim3 block(HEAT_TX, HEAT_WY);
    heat_fused_kernel<D, MODE, R><<<grid, block, 0, ctx->stream>>>(in, out, g, hk, ntx, nty, nzc,
                                                                   zc);
    B2R_LAUNCH_CHECK();
    return B2R_OK;
}

template <int MODE>
static int heat_fused_mode(b2r_ctx* ctx, const float* in, float* out, const HeatK& hk, int D)
{
    const int R = ctx->g.R;
    switch (D * 8 + R) {
    case 1 * 8 + 1: return heat_fused_t<1, MODE, 1>(ctx, in, out, hk);
    case 1 * 8 + 2: return heat_fused_t<1, MODE, 2>(ctx, in, out, hk);
    case 1 * 8 + 3: return heat_fused_t<1, MODE, 3>(ctx, in, out, hk);
    case 1 * 8 + 4: return heat_fused_t<1, MODE, 4>(ctx, in, out, hk);
    case 1 * 8 + 5: return heat_fused_t<1, MODE, 5>(ctx, in, out, hk);
    case 1 * 8 + 6: return heat_fused_t<1, MODE, 6>(ctx, in, out, hk);
    case 1 * 8 + 7: return heat_fused_t<1, MODE, 7>(ctx, in, out, hk);
    case 1 * 8 + 8: return heat_fused_t<1, MODE, 8>(ctx, in, out, hk);
    case 2 * 8 + 1: return heat_fused_t<2, MODE, 1>(ctx, in, out, hk);
    case 2 * 8 + 2: return heat_fused_t<2, MODE, 2>(ctx, in, out, hk);
    case 2 * 8 + 3: return heat_fused_t<2, MODE, 3>(ctx, in, out, hk);
    case 2 * 8 + 4: return heat_fused_t<2, MODE, 4>(ctx, in, out, hk);
    case 3 * 8 + 1: return heat_fused_t<3, MODE, 1>(ctx, in, out, hk);
    case 3 * 8 + 2: return heat_fused_t<3, MODE, 2>(ctx, in, out, hk);
    case 4 * 8 + 1: return heat_fused_t<4, MODE, 1>(ctx, in, out, hk);
    case 4 * 8 + 2: return heat_fused_t<4, MODE, 2>(ctx, in, out, hk);
    }
    b2r_set_error("fused heat needs R*D <= 8 (R=%d, D=%d)", R, D);
    return B2R_E_UNSUPPORTED;
}

int b2r_heat_fused_launch(b2r_ctx* ctx, const void* in, void* out, const b2r_heat_params* p)
{
    const HeatK hk = heat_k(p);
    if (p->mode == B2R_HEAT_F32)
        return heat_fused_mode<B2R_HEAT_F32>(ctx, (const float*)in, (float*)out, hk, p->D);
    return heat_fused_mode<B2R_HEAT_REF>(ctx, (const float*)in, (float*)out, hk, p->D);
}
