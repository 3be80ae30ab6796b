// b2r_gol.hip -- 3D Game of Life step (26-neighbour Moore count) for gfx950.
//
// Reference: gol_3d_kernel gol-3D.cu:30-102 and gol_3D_gpu_main :119-155.  The
// rule (gol-3D.cu:93-99): alive & n<=1 -> 0, dead & n==5 -> 1, alive & n>=8 -> 0,
// otherwise unchanged.  The reference leaves tile-corner halos unloaded,
// applies the rule with an uninitialised count and writes only changed cells
// into a ping-pong buffer (SURVEY defect D5); none of that is reproduced: every
// cell of the update region is written with the total rule.
//
// SWAR formulation: a thread owns one aligned dword = 4 cells along x and RY
// rows, and marches z.  With cells in {0,1} every byte-wise sum of up to 27
// cells fits in a byte, so the 3x3x3 box sum (n + self) is built with plain
// 32-bit adds: z sums per row, then y sums, then the x neighbours by
// byte-aligning the dword with a packed edge word E = [L3 .. R0] that holds the
// last cell of the left dword and the first cell of the right one.  The rule is
// evaluated on all four bytes at once with carry-free byte compares:
//   new = (box == 5) | (self & 3 <= box <= 8)
// (box = n + self: dead cells are born at n == 5, live ones survive 2 <= n <= 7).
#include "b2r_common.h"

__device__ __forceinline__ uint32_t gol_rule4(uint32_t box, uint32_t cur)
{
    const uint32_t H = 0x80808080u;
    const uint32_t ge3 = (box + 0x7D7D7D7Du) & H;
    const uint32_t ge5 = (box + 0x7B7B7B7Bu) & H;
    const uint32_t ge6 = (box + 0x7A7A7A7Au) & H;
    const uint32_t ge9 = (box + 0x77777777u) & H;
    const uint32_t born = ge5 & ~ge6;
    const uint32_t keep = (ge3 & ~ge9) & (cur << 7);
    return (born | keep) >> 7;
}

constexpr int GOL_TXD = 32;  // dword lanes per row
constexpr int GOL_WY = 8;    // thread rows per workgroup

// Load the aligned dword starting at x = xf of a row (0 if the dword does not
// touch [0, S_x) or the row/plane is out of the block).
__device__ __forceinline__ uint32_t gol_ld(const uint8_t* __restrict__ row, int xf, int Sx, bool ok)
{
    if (!ok || xf + 3 < 0 || xf > Sx - 1) return 0u;
    return *reinterpret_cast<const uint32_t*>(row + xf);
}

template <int RY>
__global__ __launch_bounds__(256) void gol_step_kernel(const uint8_t* __restrict__ in,
                                                       uint8_t* __restrict__ out, KGeo g,
                                                       int b2r_i, int jd0, int ntx, int nty,
                                                       int zc)
{
    constexpr int TY = GOL_WY * RY;
    const int lane = threadIdx.x, wy = threadIdx.y;
    int t = blockIdx.x;
    const int txi = t % ntx;
    t /= ntx;
    const int tyi = t % nty;
    const int zci = t / nty;
    const int b = blockIdx.y;

    const Box box = update_box(g, b, 1, b2r_i);
    const int jd = jd0 + txi * GOL_TXD + lane;
    const int xf = g.h + 4 * jd;  // first cell of this thread's dword
    const int y0 = 1 + tyi * TY + wy * RY;
    const int zs = 1 + zci * zc;
    const int zlo = max(zs, box.lo[2]);
    const int zhi = min(min(zs + zc, g.S[2] - 1) - 1, box.hi[2]);
    if (zlo > zhi) return;  // no barriers in this kernel: per-thread exits are fine

    // byte mask of cells inside the x update range
    uint32_t xm = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (xf + k >= box.lo[0] && xf + k <= box.hi[0]) xm |= 0xFFu << (8 * k);
    if (xm == 0) return;

    const long long boff = (long long)b * g.bstride + g.xoff;
    const uint8_t* __restrict__ src = in + boff;
    uint8_t* __restrict__ dst = out + boff;
    const long long slice = g.slice;
    const int pitch = g.pitch;
    const int Sx = g.S[0];

    // rows y0-1 .. y0+RY, planes z-1, z, z+1
    uint32_t c[RY + 2][3], e[RY + 2][3];
    bool rok[RY + 2];
#pragma unroll
    for (int r = 0; r < RY + 2; ++r) {
        const int y = y0 - 1 + r;
        rok[r] = y >= 0 && y < g.S[1];
    }
    auto load_plane = [&](int z, int slot) {
        const long long pz = (long long)z * slice;
#pragma unroll
        for (int r = 0; r < RY + 2; ++r) {
            const uint8_t* row = src + pz + (long long)(y0 - 1 + r) * pitch;
            const uint32_t C = gol_ld(row, xf, Sx, rok[r]);
            const uint32_t L = gol_ld(row, xf - 4, Sx, rok[r]);
            const uint32_t Rr = gol_ld(row, xf + 4, Sx, rok[r]);
            c[r][slot] = C;
            e[r][slot] = (L & 0xFF000000u) | (Rr & 0x000000FFu);
        }
    };
    load_plane(zlo - 1, 0);
    load_plane(zlo, 1);

    for (int z = zlo; z <= zhi; ++z) {
        load_plane(z + 1, 2);
        uint32_t zc_[RY + 2], ze_[RY + 2];
#pragma unroll
        for (int r = 0; r < RY + 2; ++r) {
            zc_[r] = c[r][0] + c[r][1] + c[r][2];
            ze_[r] = e[r][0] + e[r][1] + e[r][2];
        }
        const long long pz = (long long)z * slice;
#pragma unroll
        for (int j = 0; j < RY; ++j) {
            const int y = y0 + j;
            if (y >= box.lo[1] && y <= box.hi[1]) {
                const uint32_t pc = zc_[j] + zc_[j + 1] + zc_[j + 2];
                const uint32_t pe = ze_[j] + ze_[j + 1] + ze_[j + 2];
                const uint32_t left = __builtin_amdgcn_alignbyte(pc, pe, 3);   // x-1
                const uint32_t right = __builtin_amdgcn_alignbyte(pe, pc, 1);  // x+1
                const uint32_t boxsum = left + pc + right;
                const uint32_t cur = c[j + 1][1];
                const uint32_t nw = gol_rule4(boxsum, cur);
                *reinterpret_cast<uint32_t*>(dst + pz + (long long)y * pitch + xf) =
                    (nw & xm) | (cur & ~xm);
            }
        }
#pragma unroll
        for (int r = 0; r < RY + 2; ++r) {
            c[r][0] = c[r][1];
            c[r][1] = c[r][2];
            e[r][0] = e[r][1];
            e[r][1] = e[r][2];
        }
    }
}

static int gol_check(b2r_ctx* ctx)
{
    B2R_CHECK(ctx, B2R_E_ARG, "null context");
    B2R_CHECK(ctx->g.elem_size == 1, B2R_E_UNSUPPORTED, "Game of Life needs u8 cells");
    B2R_CHECK(ctx->g.radius == 1, B2R_E_ARG, "Game of Life needs radius 1 (got %d)",
              ctx->g.radius);
    return B2R_OK;
}

int b2r_gol_launch(b2r_ctx* ctx, const void* in, void* out, int i)
{
    constexpr int RY = 4;
    constexpr int TY = GOL_WY * RY;
    const KGeo& g = ctx->k;
    // dwords covering the widest update range [1, S_x-2] (floor division by 4)
    const int jd0 = (1 - g.h) >> 2;
    const int jd1 = (g.S[0] - 2 - g.h) >> 2;
    const int ntx = ceil_div(jd1 - jd0 + 1, GOL_TXD);
    const int nty = ceil_div(g.S[1] - 2, TY);
    const int nz = g.S[2] - 2;
    long long tiles = (long long)ntx * nty * g.nblk;
    int nzc = (int)((4096 + tiles - 1) / tiles);
    int zc = ceil_div(nz, nzc < 1 ? 1 : nzc);
    if (zc < 8) zc = 8;
    if (zc > nz) zc = nz;
    nzc = ceil_div(nz, zc);
    dim3 grid(ntx * nty * nzc, g.nblk);
    dim3 block(GOL_TXD, GOL_WY);
    gol_step_kernel<RY><<<grid, block, 0, ctx->stream>>>((const uint8_t*)in, (uint8_t*)out, g, i,
                                                         jd0, ntx, nty, zc);
    B2R_LAUNCH_CHECK();
    return B2R_OK;
}

extern "C" int b2r_gol_step(b2r_ctx* ctx, const void* d_in, void* d_out, int b2r_i)
{
    if (int rc = gol_check(ctx)) return rc;
    if (int rc = b2r_check_field(ctx, d_in)) return rc;
    if (int rc = b2r_check_field(ctx, d_out)) return rc;
    B2R_CHECK(d_in != d_out, B2R_E_ARG, "in and out must differ");
    B2R_CHECK(b2r_i >= 0 && b2r_i < ctx->g.R, B2R_E_ARG, "b2r_i=%d outside [0,R)", b2r_i);
    if (int rc = b2r_activate(ctx)) return rc;
    return b2r_gol_launch(ctx, d_in, d_out, b2r_i);
}

extern "C" int b2r_gol_run(b2r_ctx* ctx, void* d_a, void* d_b, int nsteps, int fused,
                           int* result_in_b, float* ms)
{
    (void)fused;
    if (int rc = gol_check(ctx)) return rc;
    if (int rc = b2r_check_field(ctx, d_a)) return rc;
    if (int rc = b2r_check_field(ctx, d_b)) return rc;
    B2R_CHECK(d_a != d_b && nsteps >= 0, B2R_E_ARG, "bad buffers or nsteps");
    if (int rc = b2r_activate(ctx)) return rc;
    if (int rc = b2r_field_copy(ctx, d_b, d_a)) return rc;  // gol-3D.cu:128-129
    void* cur = d_a;
    void* nxt = d_b;
    if (int rc = b2r_time_begin(ctx, ms)) return rc;
    const int R = ctx->g.R;
    for (int s = 0; s < nsteps; s += R) {
        if (int rc = b2r_direct_exchange(ctx, &cur, 1)) return rc;
        const int n = nsteps - s < R ? nsteps - s : R;
        for (int i = 0; i < n; ++i) {
            if (int rc = b2r_gol_launch(ctx, cur, nxt, i)) return rc;
            void* t = cur;
            cur = nxt;
            nxt = t;
        }
    }
    if (int rc = b2r_time_end(ctx, ms)) return rc;
    if (result_in_b) *result_in_b = cur == d_b;
    return B2R_OK;
}
