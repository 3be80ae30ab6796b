// b2r_gol.hip -- 3D Game of Life step (26-neighbour Moore count) for gfx950.
//
// Reference: gol_3d_kernel gol-3D.cu:30-102 and gol_3D_gpu_main :119-155.  The
// rule (gol-3D.cu:93-99): alive & n<=1 -> 0, dead & n==5 -> 1, alive & n>=8 -> 0,
// otherwise unchanged.  The reference leaves tile-corner halos unloaded,
// applies the rule with an uninitialised count and writes only changed cells
// into a ping-pong buffer; none of that is reproduced: every
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
                                                       int nzc, int zc)
{
    constexpr int TY = GOL_WY * RY;
    const int lane = threadIdx.x, wy = threadIdx.y;
    int t = xcd_tile(blockIdx.x, gridDim.x);
    const int txi = t % ntx;
    t /= ntx;
    const int tyi = t % nty;
    t /= nty;
    const int zci = t % nzc;
    const int b = t / nzc;

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

    // rows y0-1 .. y0+RY, planes z-1, z, z+1, and z+2 loaded a plane ahead
    uint32_t c[RY + 2][4], e[RY + 2][4];
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
    load_plane(zlo + 1, 2);

    for (int z = zlo; z <= zhi; ++z) {
        if (z + 2 <= zhi + 1) load_plane(z + 2, 3);
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
            c[r][2] = c[r][3];
            e[r][0] = e[r][1];
            e[r][1] = e[r][2];
            e[r][2] = e[r][3];
        }
    }
}

// ---------------------------------------------------------------------------
// R generations in one launch: temporal blocking (B2R's deep halo on chip).
// A workgroup owns an output tile of 16*TXQ cells x TY = NR-2R rows of one
// block and marches z through a chunk of planes.  The tile shape (TXQ, NR) is
// a template parameter, chosen per context by the gol_tile tuning field.  Each
// thread owns one 16-cell quad position of the tile grown by one quad per side
// in x and R rows per side in y, and keeps, for every generation t = 0..R-1, the last three
// planes of its own position in registers.  For every input plane z_in it
// computes generation t (t = 1..R) of plane p = z_in - t from generation t-1:
//   zs  = sum of its three planes of generation t-1 (registers, SWAR bytes),
//   one LDS plane of zs (double-buffered, one barrier per generation),
//   box = zs(y-1) + zs + zs(y+1), plus the neighbour quads' edge dword
//         columns for x-1 / x+1 via byte alignment, then the rule with the
//         cell's own previous value; kept in registers for generation t+1 or,
//         for generation R, written to global memory.
// The valid region shrinks by one cell per generation, which the 16-cell
// x halo and the R halo rows cover.  Cells outside a generation's update box
// (frozen global faces) keep their previous value, as in the step-by-step
// path.  Output quads start inside the owned block; when B_x is not a
// multiple of 16 the last quad of a row also writes halo cells beyond B_x:
// at a frozen face with their unchanged values, elsewhere with values the
// next exchange overwrites.
// ---------------------------------------------------------------------------
// Tile shapes: gol_tile 1, 2, 3 select this kernel with TXQ output quads (16
// cells) per tile row and NR rows including the R halo rows per side
// (GF_SHAPES); gol_tile 4-15 select the bit-sliced kernel below with NW
// output words (32 cells) per row, NR rows, RT rows per thread and XD = 1
// for waves of whole block rows, XD = 2 for waves of whole periodic grid rows
// (GB_SHAPES); 0 (automatic) is the bit-sliced 8 x 64 shape.
constexpr int GF_SHAPES[3][2] = {{8, 32}, {8, 48}, {16, 56}};
constexpr int GB_SHAPES[12][4] = {{8, 32, 1, 0},  {8, 64, 1, 0},  {8, 96, 1, 0},  {8, 40, 1, 0},
                                  {8, 72, 1, 0},  {8, 72, 2, 0},  {8, 136, 2, 0}, {8, 72, 1, 1},
                                  {8, 48, 1, 1},  {16, 72, 2, 2}, {16, 40, 1, 2}, {16, 64, 1, 2}};
constexpr int GF_PD = 2;  // input planes in flight per thread

__device__ __forceinline__ uint4 q_add(uint4 a, uint4 b)
{
    return make_uint4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

// byte mask of the four cells x .. x+3 inside [lo, hi]
__device__ __forceinline__ uint32_t dword_mask(int x, int lo, int hi)
{
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (x + k >= lo && x + k <= hi) m |= 0xFFu << (8 * k);
    return m;
}

template <int R, int TXQ, int NR>
__global__ __launch_bounds__((TXQ + 2) * NR) void gol_fused_kernel(const uint8_t* __restrict__ in,
                                                                   uint8_t* __restrict__ out,
                                                                   KGeo g, int ntx, int nty,
                                                                   int nzc, int zc)
{
    constexpr int GF_NQ = TXQ + 2;         // quads per row: plus one per side
    constexpr int GF_PAD = GF_NQ + 1;      // LDS rows read at +-(NQ+1) stay in bounds
    constexpr int TY = NR - 2 * R;
    constexpr int ZSB = GF_NQ * NR + 2 * GF_PAD;
    __shared__ uint4 zsbuf[2 * ZSB];
    // the x and w dwords of zs again, densely, so that the neighbour-column
    // reads of consecutive lanes hit consecutive banks
    __shared__ uint32_t zxbuf[2 * ZSB], zwbuf[2 * ZSB];

    const int tid = threadIdx.x;
    const int r = tid / GF_NQ, q = tid % GF_NQ;
    int tt = xcd_tile(blockIdx.x, gridDim.x);
    const int txi = tt % ntx;
    tt /= ntx;
    const int tyi = tt % nty;
    tt /= nty;
    const int zci = tt % nzc;
    const int b = tt / nzc;
    const int h = g.h;
    const int x0 = h + txi * 16 * TXQ;  // first output cell
    const int y0 = h + tyi * TY;
    const int zc0 = h + zci * zc;
    const int zc1 = min(zc0 + zc, h + g.B[2]);
    const long long boff = (long long)b * g.bstride + g.xoff;
    const uint8_t* __restrict__ src = in + boff;
    uint8_t* __restrict__ dst = out + boff;
    const long long slice = g.slice;
    const int pitch = g.pitch;

    // this thread's quad: x = x .. x+15, row y
    const int x = x0 + 16 * (q - 1);
    const int y = y0 - R + r;
    const bool ld_ok = y < g.S[1] && x + 15 >= 0 && x < g.S[0];
    // every thread loads (no branch, so a load in flight is never copied):
    // threads outside the block read the tile's first quad of the last row,
    // in bounds and aligned, and use zero instead
    const uint8_t* __restrict__ col =
        src + (ld_ok ? (long long)y * pitch + x : (long long)min(y, g.S[1] - 1) * pitch + x0);
    auto load = [&](int z) { return *reinterpret_cast<const uint4*>(col + (long long)z * slice); };
    const uint32_t ldm = ld_ok ? ~0u : 0u;

    // per-generation boxes (frozen faces) and this thread's update masks
    Box bx[R];
    bool clip[R];
#pragma unroll
    for (int t = 1; t <= R; ++t) {
        bx[t - 1] = update_box(g, b, 1, t - 1);
        const int gr = R - t;  // growth of generation t's region
        const Box& B_ = bx[t - 1];
        clip[t - 1] = x0 - 16 < B_.lo[0] || x0 + 16 * (TXQ + 1) - 1 > B_.hi[0] ||
                      y0 - gr < B_.lo[1] || y0 + TY + gr - 1 > B_.hi[1] ||
                      zc0 - gr < B_.lo[2] || zc1 + gr - 1 > B_.hi[2];
    }
    const bool out_ok = q >= 1 && q <= GF_NQ - 2 && r >= R && r < R + TY && y < h + g.B[1] &&
                        x < h + g.B[0];

    // gen[t][0..2]: generation t at planes (newest-2, newest-1, newest)
    uint4 gen[R][3];
    // input planes are loaded GF_PD planes ahead of their use into slots that
    // take turns (the loop is unrolled by GF_PD), so no register copy ever
    // waits for a load still in flight
    const int zend = zc1 + R;
    int lc = 0;  // generations computed so far (selects the LDS buffer)
    auto plane = [&](int zin, uint4& slot) {
        gen[0][0] = gen[0][1];
        gen[0][1] = gen[0][2];
        gen[0][2] = make_uint4(slot.x & ldm, slot.y & ldm, slot.z & ldm, slot.w & ldm);
        slot = load(min(zin + GF_PD, zend - 1));  // the last reloads are unused
#pragma unroll
        for (int t = 1; t <= R; ++t) {
            // uniform: generation t-1 not yet 3 planes, or a plane past the chunk
            if (zin < zc0 - R + 2 * t || zin >= zend) continue;
            const int p = zin - t;
            const uint4 zs = q_add(q_add(gen[t - 1][0], gen[t - 1][1]), gen[t - 1][2]);
            const int bo = (lc & 1) * ZSB + GF_PAD;
            uint4* buf = zsbuf + bo;
            uint32_t* bx_ = zxbuf + bo;
            uint32_t* bw_ = zwbuf + bo;
            ++lc;
            buf[tid] = zs;
            bx_[tid] = zs.x;
            bw_[tid] = zs.w;
            __syncthreads();
            const uint4 P = q_add(q_add(buf[tid - GF_NQ], zs), buf[tid + GF_NQ]);
            const int il = tid - 1, ir = tid + 1;
            const uint32_t el = bw_[il - GF_NQ] + bw_[il] + bw_[il + GF_NQ];
            const uint32_t er = bx_[ir - GF_NQ] + bx_[ir] + bx_[ir + GF_NQ];
            const uint4 self = gen[t - 1][1];
            uint4 nw;
            nw.x = gol_rule4(__builtin_amdgcn_alignbyte(P.x, el, 3) + P.x +
                             __builtin_amdgcn_alignbyte(P.y, P.x, 1), self.x);
            nw.y = gol_rule4(__builtin_amdgcn_alignbyte(P.y, P.x, 3) + P.y +
                             __builtin_amdgcn_alignbyte(P.z, P.y, 1), self.y);
            nw.z = gol_rule4(__builtin_amdgcn_alignbyte(P.z, P.y, 3) + P.z +
                             __builtin_amdgcn_alignbyte(P.w, P.z, 1), self.z);
            nw.w = gol_rule4(__builtin_amdgcn_alignbyte(P.w, P.z, 3) + P.w +
                             __builtin_amdgcn_alignbyte(er, P.w, 1), self.w);
            if (clip[t - 1]) {
                const Box& B_ = bx[t - 1];
                const bool in_yz =
                    y >= B_.lo[1] && y <= B_.hi[1] && p >= B_.lo[2] && p <= B_.hi[2];
                uint32_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
                if (in_yz) {
                    m0 = dword_mask(x, B_.lo[0], B_.hi[0]);
                    m1 = dword_mask(x + 4, B_.lo[0], B_.hi[0]);
                    m2 = dword_mask(x + 8, B_.lo[0], B_.hi[0]);
                    m3 = dword_mask(x + 12, B_.lo[0], B_.hi[0]);
                }
                nw.x = (nw.x & m0) | (self.x & ~m0);
                nw.y = (nw.y & m1) | (self.y & ~m1);
                nw.z = (nw.z & m2) | (self.z & ~m2);
                nw.w = (nw.w & m3) | (self.w & ~m3);
            }
            if (t < R) {
                gen[t][0] = gen[t][1];
                gen[t][1] = gen[t][2];
                gen[t][2] = nw;
            } else if (out_ok) {
                *reinterpret_cast<uint4*>(dst + (long long)p * slice + (long long)y * pitch + x) = nw;
            }
        }
    };
    const int z0 = zc0 - R;
    uint4 sl[GF_PD];
#pragma unroll
    for (int k = 0; k < GF_PD; ++k) sl[k] = load(min(z0 + k, zend - 1));
    // whole groups of GF_PD planes; planes past zend compute nothing
    for (int zin = z0; zin < zend; zin += GF_PD) {
#pragma unroll
        for (int k = 0; k < GF_PD; ++k) plane(zin + k, sl[k]);
    }
}

// ---------------------------------------------------------------------------
// Bit-sliced form of the fused kernel (gol_tile 0, 4..15): R generations per launch
// as above, but a thread owns a 32-cell word of a row, bit i = cell x+i,
// packed from the byte cells when a plane is loaded and unpacked when
// generation R is stored.  Per generation and plane:
//   S   = z-sum of its three planes of generation t-1, two bit planes (0..3);
//   XZ  = S(x-1) + S + S(x+1), the x neighbours' z-sums (through an LDS
//         plane, or DPP lane moves with XD) shifted in across word
//         boundaries, four bit planes (0..9);
//   box = XZ(y-1) + XZ + XZ(y+1) through an LDS plane;
//   the rule: alive next iff box == 5, or alive and 3 <= box <= 8.
// Every adder is built from full adders of two 3-input logic operations
// (xor3, majority); with XD a generation compiles to about 40 VALU
// instructions per 32 cells, against about 150 for the byte-lane form.
// With pull set (B_x a multiple of 32), a word outside the block's owned
// cells is loaded from the neighbouring block that owns it, so no exchange is
// needed before the launch (its X phase alone moved 1-4 bytes per row, one
// cache line each, a quarter of the R=4 window).  The source is resolved as
// the ordered X, Y, Z exchange would have filled that halo cell: along z
// first, then y, then x, moving to the neighbour across each face the word
// lies beyond and stopping at a frozen face (no neighbour), where the cell
// keeps its own value.  Words lie on the 16-byte grid of x = h,
// so each half of a word (16 cells) lies inside or outside the padded row as a
// whole: halo words load only their half next to the tile, a row's last word
// loads and stores its second half only when that half is in the row.  Cells
// hold 0 or 1.
// ---------------------------------------------------------------------------
// 32 byte cells (a: cells 0..15, b: 16..31) -> one word, bit i = cell i (bit
// 0 of each byte).  Byte j of dword k is cell 4k+j: a byte dot product of the
// masked dword with the weights 1, 2, 4, 8 (16, 32, 64, 128 for odd k, added
// onto the even dword's) gives the 8 bits of cells 8p..8p+7 (dwords 2p and
// 2p+1) in the low byte, and two byte permutes and a shift-or place the four
// bytes: 21 instructions against 43 for a bit-matrix transpose.
__device__ __forceinline__ uint32_t gb_pack(uint4 a, uint4 b)
{
    const uint32_t m = 0x01010101u, lo4 = 0x08040201u, hi4 = 0x80402010u;
    auto two = [&](uint32_t e, uint32_t o) {
        return __builtin_amdgcn_udot4(o & m, hi4, __builtin_amdgcn_udot4(e & m, lo4, 0u, false),
                                      false);
    };
    const uint32_t n0 = two(a.x, a.y), n1 = two(a.z, a.w), n2 = two(b.x, b.y), n3 = two(b.z, b.w);
    // v_perm_b32 byte selectors: 0-3 pick the second operand's bytes, 4-7 the
    // first's, 0x0c a zero byte: (n0.b0, n1.b0, 0, 0)
    const uint32_t lo = __builtin_amdgcn_perm(n1, n0, 0x0c0c0400u);
    const uint32_t hi = __builtin_amdgcn_perm(n3, n2, 0x0c0c0400u);
    return lo | (hi << 16);
}

// One word -> 32 byte cells, the inverse of gb_pack.  Dword k holds cells
// 4k..4k+3: their nibble of the word times 0x204081 (bits 0, 7, 14, 21) puts
// nibble bit i at bits i, i+7, i+14, i+21, 16 distinct bits for the 4 nibble
// bits, so nothing carries, and bit 8i holds nibble bit i; the byte mask keeps
// bits 0, 8, 16, 24.
__device__ __forceinline__ uint32_t gb_spread4(uint32_t w, int k)
{
    return __builtin_amdgcn_ubfe(w, 4 * k, 4) * 0x204081u & 0x01010101u;
}
__device__ __forceinline__ void gb_unpack(uint32_t w, uint4& a, uint4& b)
{
    a = make_uint4(gb_spread4(w, 0), gb_spread4(w, 1), gb_spread4(w, 2), gb_spread4(w, 3));
    b = make_uint4(gb_spread4(w, 4), gb_spread4(w, 5), gb_spread4(w, 6), gb_spread4(w, 7));
}

// The two outputs of a full adder, each one v_bitop3 (truth tables 0x96 and
// 0xE8, both symmetric in the three inputs).  Written as C expressions the
// compiler shares a ^ b between them and emits three instructions for two.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
__device__ __forceinline__ uint32_t maj3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xE8);
}

// The x-sum of a word's z-sums (s0, s1: two bit planes, 0..3) and those of
// its x neighbours (L: word x-32, Rg: word x+32): bit i of l is cell x+i-1,
// of rr cell x+i+1; XZ = l + S + rr in four bit planes (0..9).
__device__ __forceinline__ uint4 gb_xsum(uint32_t s0, uint32_t s1, uint2 L, uint2 Rg)
{
    const uint32_t l0 = __builtin_amdgcn_alignbit(s0, L.x, 31),
                   l1 = __builtin_amdgcn_alignbit(s1, L.y, 31);
    const uint32_t r0 = __builtin_amdgcn_alignbit(Rg.x, s0, 1),
                   r1 = __builtin_amdgcn_alignbit(Rg.y, s1, 1);
    const uint32_t x0b = xor3(l0, s0, r0), k1 = maj3(l0, s0, r0);
    const uint32_t u = xor3(l1, s1, r1), v = maj3(l1, s1, r1);
    const uint32_t x1b = u ^ k1, w = u & k1;
    return make_uint4(x0b, x1b, v ^ w, v & w);
}

// The rule on box = U + X + Dn (x-sums of rows y-1, y, y+1, each <= 9): alive
// next iff box == 5, or alive and 3 <= box <= 8.  The sum leaves bits b0..b3
// and the two carries out of bit 3, c4a and c4b (box >= 16 iff either is
// set; box < 32, so never both).  Below 16, with b3 set box is 8..15 and the
// cell lives on iff it is alive and box == 8 (b2 b1 b0 = 0); with b3 clear
// box is 0..7 and the cell is alive next iff box == 5, or it is alive and
// box >= 3 (checked over all inputs 0..9 and both cell states).
__device__ __forceinline__ uint32_t gb_rule(uint4 U, uint4 X, uint4 Dn, uint32_t self)
{
    const uint32_t b0 = xor3(U.x, X.x, Dn.x), c1 = maj3(U.x, X.x, Dn.x);
    const uint32_t e1 = xor3(U.y, X.y, Dn.y), c2a = maj3(U.y, X.y, Dn.y);
    const uint32_t b1 = e1 ^ c1, c2b = e1 & c1;
    const uint32_t e2 = xor3(U.z, X.z, Dn.z), c3a = maj3(U.z, X.z, Dn.z);
    const uint32_t b2 = xor3(e2, c2a, c2b), c3b = maj3(e2, c2a, c2b);
    const uint32_t e3 = xor3(U.w, X.w, Dn.w), c4a = maj3(U.w, X.w, Dn.w);
    const uint32_t b3 = xor3(e3, c3a, c3b), c4b = maj3(e3, c3a, c3b);
    const uint32_t lo = b2 | b1 | b0;
    const uint32_t ge3 = b2 | (b1 & b0);
    const uint32_t eq5 = b2 & ~b1 & b0;
    const uint32_t m = eq5 | (self & ge3);
    const uint32_t z = b3 & self & ~lo;
    const uint32_t tt = z | (~b3 & m);
    return tt & ~c4a & ~c4b;
}

// Whole periodic grid rows of 16 words, one per 16-lane DPP row: the z-sums
// of words q-1 and q+1 of the row are those of lanes (q-1) mod 16 and
// (q+1) mod 16 of the DPP row (row_ror:1, row_ror:15).
__device__ __forceinline__ uint2 gb_row_prev(uint32_t s0, uint32_t s1)
{
    return make_uint2(__builtin_amdgcn_mov_dpp(s0, 0x121, 0xf, 0xf, true),
                      __builtin_amdgcn_mov_dpp(s1, 0x121, 0xf, 0xf, true));
}
__device__ __forceinline__ uint2 gb_row_next(uint32_t s0, uint32_t s1)
{
    return make_uint2(__builtin_amdgcn_mov_dpp(s0, 0x12F, 0xf, 0xf, true),
                      __builtin_amdgcn_mov_dpp(s1, 0x12F, 0xf, 0xf, true));
}

// bits of the cells x .. x+31 inside [lo, hi]
__device__ __forceinline__ uint32_t word_mask(int x, int lo, int hi)
{
    const int a = max(lo - x, 0), e = min(hi - x, 31);
    if (a > e) return 0u;
    const uint32_t upto = e >= 31 ? ~0u : (2u << e) - 1u;
    return upto & ~((1u << a) - 1u);
}

// Packed fields: a block's row is PW = B_x/32 + 2 words, word w holding
// cells x = h-32+32w .. h+32w-1 (bit i = cell x+i), planes of S_y rows, blocks
// of S_z planes; the kernel then loads and stores words instead of packing
// and unpacking byte cells.  PGeo carries the byte strides of the packed form.
// IO bit 0: the input is packed (else byte cells); bit 1: the output is
// packed.  IO 3 runs between packed fields, IO 0 between byte fields; IO 2
// packs the byte input as it loads it (a run's first window) and IO 1
// unpacks as it stores (its last), saving the separate pack and unpack passes.
struct PGeo {
    long long rb, sb, bb;  // bytes per row, plane and block
    int pw;                // words per packed row
};

// threads of a bit-sliced workgroup: one per word and RT rows, or with XD
// whole waves of 64 / (NW + 2) rows each (XD = 2: 64 / NW rows, no halo words)
__host__ __device__ constexpr int gb_threads(int NW, int NR, int RT, int XD)
{
    return XD ? NR / RT / (64 / (XD == 2 ? NW : NW + 2)) * 64 : (NW + 2) * NR / RT;
}

// XD = 1: every wave holds whole rows (64 / NQ rows, the lanes past them
// idle), so a word's x neighbours are the next lanes: the z-sums reach them by
// wave-wide DPP lane shifts instead of a trip through LDS and a barrier, and
// the x-sums alternate between two LDS buffers, so one barrier per
// generation and plane remains instead of two.  A row's first and last words
// are halo words; the shift brings them another row's word (or nothing at the
// wave's ends), which only touches bits that no output cell reads within R <=
// 4 generations, as with the LDS path's spare words.
// XD = 2 (gol_tile 13-15): the grid is periodic in x over 16 words (512
// cells, the x blocks side by side), and every wave holds 4 whole grid rows,
// one per 16-lane DPP row, lane q of a row holding word q of the grid row
// (block q / (B_x/32)).  A word's x neighbours, the periodic wrap included,
// come from a rotation within the DPP row (row_ror), so no lane holds a halo
// word: all 64 lanes own output, against 48 of 64 with XD = 1.
template <int R, int NW, int NR, int IO = 0, int RT = 1, int XD = 0>
__global__ __launch_bounds__(gb_threads(NW, NR, RT, XD)) void gol_bits_kernel(
    const uint8_t* __restrict__ in, uint8_t* __restrict__ out, KGeo g, int ntx, int nty, int nzc,
    int zc, int pull, PGeo pg)
{
    constexpr bool PKI = (IO & 1) != 0, PKO = (IO & 2) != 0;
    constexpr bool WR = XD == 2;          // whole periodic grid rows
    constexpr int NQ = WR ? NW : NW + 2;  // words per row: plus one per side
    constexpr int NT = NQ * NR;
    constexpr int TY = NR - 2 * R;
    // a thread owns RT rows of the tile, NR/RT apart: twice the work between
    // the two barriers of a generation at RT = 2, for half as many waves
    static_assert(NR % RT == 0, "rows per thread must divide the tile rows");
    constexpr int RW = 64 / NQ;  // rows per wave with XD
    static_assert(!XD || ((WR || RT == 1) && (NR / RT) % RW == 0),
                  "XD: whole waves (XD = 1: one row per thread)");
    static_assert(!WR || NW == 16, "XD = 2: a grid row is one 16-lane DPP row");
    constexpr int NTH = NT / RT;  // (word, row) slots per row set
    __shared__ uint2 sb[XD ? 1 : NT + 2];        // z-sums of a plane, a spare word before and after
    __shared__ uint4 xb[XD ? 2 : 1][NT + 2 * NQ];  // x-sums, a spare row above and below

    const int tid = threadIdx.x;
    // this thread's word q of its row r0 (of row set 0); lanes past a wave's
    // rows are idle: they load nothing, store nothing and write no x-sums
    int q, r0;
    bool act = true;
    if (WR) {
        q = tid & 15;
        r0 = tid >> 4;
    } else if (XD) {
        const int wv = tid >> 6, l = tid & 63;
        act = l < RW * NQ;
        q = act ? l % NQ : 0;
        r0 = wv * RW + (act ? l / NQ : RW - 1);
    } else {
        q = tid % NQ;
        r0 = tid / NQ;
    }
    // workgroup order: x tile, then the block's x coordinate, then y tile, z
    // chunk and the block's y, z coordinates, so the tiles of x-neighbouring
    // blocks over the same rows run together on one XCD and a pulled halo word
    // is found in the L2 their owner filled (counting y tiles and z chunks
    // across blocks as well, so that y and z neighbours across a block face
    // also run together, was measured slower: 4589 against 4727 GLUP/s at R=4).
    // With XD = 2 one tile spans the x blocks and a lane's word q picks its
    // block, q / (B_x/32), where B_x/32 = 16/nb_x is a power of two.
    int tt = xcd_tile(blockIdx.x, gridDim.x);
    const int lwb = WR ? __builtin_ctz((unsigned)(g.B[0] >> 5)) : 0;
    int txi = 0, bxi;
    if (WR) {
        bxi = q >> lwb;
    } else {
        txi = tt % ntx;
        tt /= ntx;
        bxi = tt % g.nb[0];
        tt /= g.nb[0];
    }
    const int tyi = tt % nty;
    tt /= nty;
    const int zci = tt % nzc;
    const int b = tt / nzc * g.nb[0] + bxi;
    // update boxes along y and z are those of every block of the tile's y, z
    // coordinates; along x (periodic with XD = 2) all blocks' boxes agree
    const int bu = WR ? tt / nzc * g.nb[0] : b;
    const int h = g.h;
    const int x0 = h + txi * 32 * NW;  // first output cell (XD = 2: of the lane's block)
    const int y0 = h + tyi * TY;
    const int zc0 = h + zci * zc;
    const int zc1 = min(zc0 + zc, h + g.B[2]);
    // byte strides and the byte offset of the word at cell xx within a row,
    // of the input and of the output form
    const long long pitch = PKI ? pg.rb : g.pitch, slice = PKI ? pg.sb : g.slice;
    const long long bstr = PKI ? pg.bb : g.bstride, xo = PKI ? 0 : g.xoff;
    auto xbyte = [&](int xx) -> long long { return PKI ? 4ll * ((xx - h) / 32 + 1) : xx; };
    const long long pitch_o = PKO ? pg.rb : g.pitch, slice_o = PKO ? pg.sb : g.slice;
    const long long boff_o = (long long)b * (PKO ? pg.bb : g.bstride) + (PKO ? 0 : g.xoff);
    const uint8_t* __restrict__ src = in + (long long)b * bstr + xo;
    uint8_t* __restrict__ dst = out + boff_o;

    // this thread's words: x = x .. x+31, rows y[j]
    const int x = WR ? x0 + 32 * (q & ((1 << lwb) - 1)) : x0 + 32 * (q - 1);
    int y[RT];
    bool ld_ok[RT];
    const uint8_t* __restrict__ col[RT];
    const uint8_t* __restrict__ colown[RT];
#pragma unroll
    for (int j = 0; j < RT; ++j) {
        y[j] = y0 - R + r0 + j * (NR / RT);
        ld_ok[j] = act && (PKI ? y[j] < g.S[1] && (x - h) / 32 + 1 < pg.pw
                               : y[j] < g.S[1] && x + 31 >= 0 && x < g.S[0]);
        // every thread loads (no branch): threads outside the block read the
        // tile's first word of the last row, in bounds and aligned, and use zero
        col[j] = src + (ld_ok[j] ? (long long)y[j] * pitch + xbyte(x)
                                 : (long long)min(y[j], g.S[1] - 1) * pitch + xbyte(x0));
        colown[j] = col[j];
    }
    // pull: col is the word's source for planes of the owned z range (y, then
    // x resolved); planes of the z halo add a uniform offset into the z
    // neighbour (the y, x resolution is the same there), or read the own block
    // (colown) at a frozen z face
    long long zoff[2] = {0, 0};
    bool znb[2] = {false, false};
    if (WR && pull) {
        // a lane's block is the tile's first block bu plus its x index, and so
        // are its y and z neighbours (neighbours across y and z keep x): the
        // neighbour indices are those of bu, uniform over the workgroup
        const int ny[2] = {neighbour_block(g, bu, 1, 0), neighbour_block(g, bu, 1, 1)};
#pragma unroll
        for (int j = 0; j < RT; ++j) {
            if (!ld_ok[j]) continue;
            const int oy = y[j] < h ? -1 : (y[j] >= h + g.B[1] ? 1 : 0);
            const int nb = ny[oy > 0];
            if (oy != 0 && nb >= 0)
                col[j] = in + (long long)(nb + bxi) * bstr + xo +
                         (long long)(y[j] - oy * g.B[1]) * pitch + xbyte(x);
        }
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const int nb = neighbour_block(g, bu, 2, side);
            znb[side] = nb >= 0;
            zoff[side] = (long long)(nb - bu) * bstr + (side ? -1ll : 1ll) * g.B[2] * slice;
        }
    } else if (pull) {
#pragma unroll
        for (int j = 0; j < RT; ++j) {
            if (!ld_ok[j]) continue;
            int bb = b, yy = y[j], xx = x;
            const int oy = y[j] < h ? -1 : (y[j] >= h + g.B[1] ? 1 : 0);
            bool go = true;
            if (oy != 0) {
                const int nb = neighbour_block(g, bb, 1, oy > 0);
                if (nb >= 0) {
                    bb = nb;
                    yy = y[j] - oy * g.B[1];
                } else {
                    go = false;
                }
            }
            const int ox = x + 32 <= h ? -1 : (x >= h + g.B[0] ? 1 : 0);
            if (go && ox != 0) {
                const int nb = neighbour_block(g, bb, 0, ox > 0);
                if (nb >= 0) {
                    bb = nb;
                    xx = x - ox * g.B[0];
                }
            }
            col[j] = in + (long long)bb * bstr + xo + (long long)yy * pitch + xbyte(xx);
        }
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const int nb = neighbour_block(g, b, 2, side);
            znb[side] = nb >= 0;
            zoff[side] = (long long)(nb - b) * bstr + (side ? -1ll : 1ll) * g.B[2] * slice;
        }
    }
    // the source column of row j at plane p (p uniform over the workgroup)
    auto colp = [&](int j, int p) -> const uint8_t* {
        if (pull && p < h) return znb[0] ? col[j] + zoff[0] : colown[j];
        if (pull && p >= h + g.B[2]) return znb[1] ? col[j] + zoff[1] : colown[j];
        return col[j];
    };
    // a halo word needs only its 16 cells next to the tile (R <= 4), and the
    // second half of a row's last word may lie past the padded row end: such a
    // half reloads the word's other half (no extra traffic) and is masked off
    const bool hi_in = x + 32 <= g.pitch - g.xoff;  // second half inside the byte row
    uint32_t ldm[RT], ma[RT], mb[RT], pm[RT];
    int oa[RT], ob[RT];
    bool out_ok[RT];
    // (XD = 2 has no halo words: every word loads both halves)
    const bool hw0 = !WR && q == 0, hw1 = !WR && q == NQ - 1;
#pragma unroll
    for (int j = 0; j < RT; ++j) {
        ldm[j] = ld_ok[j] ? ~0u : 0u;
        oa[j] = hw0 && ld_ok[j] ? 16 : 0;  // the clamped word reads its first half
        const bool b_ld = ld_ok[j] && !hw1 && hi_in;
        ob[j] = b_ld ? 16 : oa[j];
        ma[j] = hw0 ? 0u : ~0u;
        mb[j] = b_ld ? ~0u : 0u;
        // gb_pack forms bits 0-15 from the first 16 bytes and bits 16-31 from
        // the second, so masking the packed word equals masking the bytes
        pm[j] = ldm[j] & ((ma[j] & 0xFFFFu) | (mb[j] & 0xFFFF0000u));
        const int r = y[j] - (y0 - R);
        out_ok[j] = act && !hw0 && !hw1 && r >= R && r < R + TY && y[j] < h + g.B[1] &&
                    x < h + g.B[0];
    }

    Box bx[R];
    bool clip[R];
    uint32_t xm[R];  // bits inside generation t's box along x
#pragma unroll
    for (int t = 1; t <= R; ++t) {
        bx[t - 1] = update_box(g, bu, 1, t - 1);
        const int gr = R - t;
        const Box& B_ = bx[t - 1];
        // (XD = 2: the tile's words are the blocks' owned words, inside every
        // box along the periodic x)
        clip[t - 1] = (!WR && (x0 - 32 < B_.lo[0] || x0 + 32 * (NW + 1) - 1 > B_.hi[0])) ||
                      y0 - gr < B_.lo[1] || y0 + TY + gr - 1 > B_.hi[1] ||
                      zc0 - gr < B_.lo[2] || zc1 + gr - 1 > B_.hi[2];
        xm[t - 1] = word_mask(x, B_.lo[0], B_.hi[0]);
    }

    uint32_t gen[RT][R][3];  // generation t at planes (newest-2, newest-1, newest)
    int xp = 0;              // XD: the x-sum buffer of the next generation step
    const int zend = zc1 + R;
    // Odd chunks march down, even chunks up, so two neighbouring chunks reach
    // their shared halo planes at the same time and the second read of them
    // hits in L2.  Input plane k of the march is zp(k); generation t is
    // computed t planes behind it.
    const int nin = zend - (zc0 - R);  // input planes of the chunk
    const int dz = (zci & 1) ? -1 : 1;
    const int zfirst = dz > 0 ? zc0 - R : zend - 1;
    auto zp = [&](int k) { return zfirst + dz * k; };
    auto plane = [&](int k, uint4 (&ld)[RT][2]) {
#pragma unroll
        for (int j = 0; j < RT; ++j) {
            uint4& la = ld[j][0];
            uint4& lb = ld[j][1];
            gen[j][0][0] = gen[j][0][1];
            gen[j][0][1] = gen[j][0][2];
            if (PKI)
                gen[j][0][2] = la.x & ldm[j];
            else
                gen[j][0][2] = gb_pack(la, lb) & pm[j];
            // the load GF_PD planes ahead (the last ones are unused)
            const int zz = zp(min(k + GF_PD, nin - 1));
            const uint8_t* pp = colp(j, zz) + (long long)zz * slice;
            if (PKI) {
                la.x = *reinterpret_cast<const uint32_t*>(pp);
            } else {
                la = *reinterpret_cast<const uint4*>(pp + oa[j]);
                lb = *reinterpret_cast<const uint4*>(pp + ob[j]);
            }
        }
        // step t (plane p) of row j from its new value: the cell keeps its
        // value outside the generation's box, then the value joins generation
        // t's planes or, for generation R, is stored
        auto finish = [&](int t, int j, int p, uint32_t nw, uint32_t self) {
            if (clip[t - 1]) {
                const Box& B_ = bx[t - 1];
                const bool in_yz =
                    y[j] >= B_.lo[1] && y[j] <= B_.hi[1] && p >= B_.lo[2] && p <= B_.hi[2];
                const uint32_t m = in_yz ? xm[t - 1] : 0u;
                nw = (nw & m) | (self & ~m);
            }
            if (t < R) {
                gen[j][t][0] = gen[j][t][1];
                gen[j][t][1] = gen[j][t][2];
                gen[j][t][2] = nw;
            } else if (out_ok[j]) {
                if (PKO) {
                    *reinterpret_cast<uint32_t*>(dst + (long long)p * slice_o +
                                                 (long long)y[j] * pitch_o +
                                                 4ll * ((x - h) / 32 + 1)) = nw;
                } else {
                    uint4 lo, hi;
                    gb_unpack(nw, lo, hi);
                    uint4* o = reinterpret_cast<uint4*>(dst + (long long)p * slice_o +
                                                        (long long)y[j] * pitch_o + x);
                    o[0] = lo;
                    if (hi_in) o[1] = hi;
                }
            }
        };
#pragma unroll
        for (int t = 1; t <= R; ++t) {
            // uniform: generation t-1 not yet 3 planes, or a plane past the chunk
            if (k < 2 * t || k >= nin) continue;
            const int p = zp(k) - dz * t;
            uint32_t s0[RT], s1[RT];
#pragma unroll
            for (int j = 0; j < RT; ++j) {
                const uint32_t ga = gen[j][t - 1][0], self = gen[j][t - 1][1], gc = gen[j][t - 1][2];
                s0[j] = xor3(ga, self, gc);
                s1[j] = maj3(ga, self, gc);
                if (!XD) sb[tid + j * NTH + 1] = make_uint2(s0[j], s1[j]);
            }
            if (!XD) __syncthreads();
            uint4* const xbp = xb[XD ? xp : 0];
            xp ^= 1;
            uint4 X[RT];
#pragma unroll
            for (int j = 0; j < RT; ++j) {
                const int pos = r0 * NQ + q + j * NTH;
                uint2 L, Rg;
                if (WR) {
                    L = gb_row_prev(s0[j], s1[j]);
                    Rg = gb_row_next(s0[j], s1[j]);
                } else if (XD) {
                    // words q-1 and q+1 of the row are lanes l-1 and l+1
                    // (bound_ctrl: the wave's first and last lanes, halo
                    // words, read 0, so no old value is set up)
                    L = make_uint2(__builtin_amdgcn_mov_dpp(s0[j], 0x138, 0xf, 0xf, true),
                                   __builtin_amdgcn_mov_dpp(s1[j], 0x138, 0xf, 0xf, true));
                    Rg = make_uint2(__builtin_amdgcn_mov_dpp(s0[j], 0x130, 0xf, 0xf, true),
                                    __builtin_amdgcn_mov_dpp(s1[j], 0x130, 0xf, 0xf, true));
                } else {
                    L = sb[pos];
                    Rg = sb[pos + 2];
                }
                X[j] = gb_xsum(s0[j], s1[j], L, Rg);
                if (act) xbp[pos + NQ] = X[j];
            }
            __syncthreads();
            // XD: a wave past the step's barrier (y-sum and rule) issues at a
            // lower priority than one still on its way to the next barrier
            // (z-sum, x-sum, LDS write), so the barriers release sooner
            // (3-5 % per R=4 launch)
            if (XD) __builtin_amdgcn_s_setprio(0);
#pragma unroll
            for (int j = 0; j < RT; ++j) {
                const int pos = r0 * NQ + q + j * NTH;
                const uint32_t self = gen[j][t - 1][1];
                finish(t, j, p, gb_rule(xbp[pos], X[j], xbp[pos + 2 * NQ], self), self);
            }
            if (XD) __builtin_amdgcn_s_setprio(1);
        }
    };
    uint4 sl[GF_PD][RT][2] = {};
#pragma unroll
    for (int k = 0; k < GF_PD; ++k) {
        const int zz = zp(min(k, nin - 1));
#pragma unroll
        for (int j = 0; j < RT; ++j) {
            const uint8_t* pp = colp(j, zz) + (long long)zz * slice;
            if (PKI) {
                sl[k][j][0].x = *reinterpret_cast<const uint32_t*>(pp);
            } else {
                sl[k][j][0] = *reinterpret_cast<const uint4*>(pp + oa[j]);
                sl[k][j][1] = *reinterpret_cast<const uint4*>(pp + ob[j]);
            }
        }
    }
    // whole groups of GF_PD planes; planes past the chunk compute nothing
    for (int k = 0; k < nin; k += GF_PD) {
#pragma unroll
        for (int jp = 0; jp < GF_PD; ++jp) plane(k + jp, sl[jp]);
    }
}

// Byte field -> two packed fields: every word of every row, halo included;
// cells outside the block (x < 0 or x >= S_x) are 0.
__global__ __launch_bounds__(256) void gol_pack_kernel(const uint8_t* __restrict__ in,
                                                       uint32_t* __restrict__ out0,
                                                       uint32_t* __restrict__ out1, KGeo g, int pw)
{
    const long long rows = (long long)g.nblk * g.S[2] * g.S[1];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * pw) return;
    const int w = (int)(i % pw);
    const long long row = i / pw;  // (block, z, y)
    const int y = (int)(row % g.S[1]);
    const long long bz = row / g.S[1];
    const int z = (int)(bz % g.S[2]);
    const int b = (int)(bz / g.S[2]);
    const int x = g.h - 32 + 32 * w;
    const uint8_t* rp = in + (long long)b * g.bstride + g.xoff + (long long)z * g.slice +
                        (long long)y * g.pitch;
    // halves of 16 cells at 16-byte aligned row offsets (xoff + h is a
    // multiple of 16), each loaded where it lies inside the padded row
    const int ra = g.xoff + x;
    uint4 a = make_uint4(0, 0, 0, 0), c = make_uint4(0, 0, 0, 0);
    if (ra >= 0 && ra + 16 <= g.pitch) a = *reinterpret_cast<const uint4*>(rp + x);
    if (ra + 16 >= 0 && ra + 32 <= g.pitch) c = *reinterpret_cast<const uint4*>(rp + x + 16);
    const uint32_t word = gb_pack(a, c) & word_mask(x, 0, g.S[0] - 1);
    out0[i] = word;
    out1[i] = word;
}


// The bit-sliced kernel reads its halo from the neighbouring blocks (pull
// above) when the block's owned x range is a whole number of 32-cell words.
static int gol_bits_pull(const KGeo& g) { return g.B[0] % 32 == 0; }

// The whole-row tiles (gol_tile 13-15, XD = 2) need a grid periodic in x over
// 16 words: B_x a multiple of 32 and nb_x * B_x = 512.  Elsewhere those tiles
// run as tile 11.
static bool gol_whole_rows(const KGeo& g)
{
    return (g.periodic & 1) && g.B[0] % 32 == 0 && g.nb[0] * g.B[0] == 512;
}

int b2r_gol_tile_ok(const b2r_ctx* ctx, int tile)
{
    return tile < 13 || gol_whole_rows(ctx->k);
}


int b2r_gol_fused_reads_halo(const b2r_ctx* ctx)
{
    const int t = ctx->tune.gol_tile;
    return (t == 0 || t >= 4) && gol_bits_pull(ctx->k) && !ctx->tune.gol_exchange;
}

static bool gol_fused_ok(const b2r_ctx* ctx, int R)
{
    return R >= 1 && R <= 4 && closed_faces(ctx->k);
}

// The z chunk of a fused launch with `tiles` tiles per plane set: the tuned
// gol_zc, or where that is 0 about four workgroup waves over 256 CUs (two
// resident per CU), at least 4R planes; at most the block's depth.
static int gol_chunk(const b2r_ctx* ctx, long long tiles, int R)
{
    const KGeo& g = ctx->k;
    int zc = ctx->tune.gol_zc;
    if (zc <= 0) {
        int nzc = (int)((2048 + tiles - 1) / tiles);
        zc = ceil_div(g.B[2], nzc < 1 ? 1 : nzc);
        if (zc < 4 * R) zc = 4 * R;
    }
    return zc > g.B[2] ? g.B[2] : zc;
}

template <int R, int TXQ, int NR>
static int gol_fused_launch_t(b2r_ctx* ctx, const void* in, void* out)
{
    constexpr int TY = NR - 2 * R;
    const KGeo& g = ctx->k;
    const int ntx = ceil_div(g.B[0], 16 * TXQ);
    const int nty = ceil_div(g.B[1], TY);
    const int zc = gol_chunk(ctx, (long long)ntx * nty * g.nblk, R);
    const int nzc = ceil_div(g.B[2], zc);
    dim3 grid(ntx * nty * nzc * g.nblk);
    gol_fused_kernel<R, TXQ, NR><<<grid, (TXQ + 2) * NR, 0, ctx->stream>>>(
        (const uint8_t*)in, (uint8_t*)out, g, ntx, nty, nzc, zc);
    B2R_LAUNCH_CHECK();
    return B2R_OK;
}

int b2r_gol_fused_chunk(const b2r_ctx* ctx)
{
    const KGeo& g = ctx->k;
    const int R = ctx->g.R;
    int tile = ctx->tune.gol_tile;
    if (tile >= 13 && !gol_whole_rows(g)) tile = 11;
    int nw, nr, wr = 0;
    if (tile >= 1 && tile <= 3) {
        nw = GF_SHAPES[tile - 1][0] / 2;  // 16-cell quads -> 32-cell words
        nr = GF_SHAPES[tile - 1][1];
    } else {
        static const int idx[16] = {1, 1, 1, 1, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
        const int* s = GB_SHAPES[idx[tile]];
        nw = s[0];
        nr = s[1];
        wr = s[3] == 2;
    }
    const int ntx = wr ? 1 : ceil_div(g.B[0], 32 * nw);
    const int nty = ceil_div(g.B[1], nr - 2 * R);
    return gol_chunk(ctx, (long long)ntx * nty * (wr ? g.nblk / g.nb[0] : g.nblk), R);
}

// words per packed row of a block: B_x/32 owned words and one per side
static int gol_pk_words(const KGeo& g) { return g.B[0] / 32 + 2; }

static PGeo gol_pgeo(const KGeo& g)
{
    PGeo pg;
    pg.pw = gol_pk_words(g);
    pg.rb = 4ll * pg.pw;
    pg.sb = pg.rb * g.S[1];
    pg.bb = pg.sb * g.S[2];
    return pg;
}

template <int R, int NW, int NR, int RT, int XD, int IO>
static int gol_bits_launch_t(b2r_ctx* ctx, const void* in, void* out)
{
    constexpr int TY = NR - 2 * R;
    const KGeo& g = ctx->k;
    // with gol_exchange the run exchanges first and the kernel reads its own halo
    const int pull = gol_bits_pull(g) && !ctx->tune.gol_exchange;
    // XD = 2: one tile spans the x blocks of its rows
    const int ntx = XD == 2 ? 1 : ceil_div(g.B[0], 32 * NW);
    const int nty = ceil_div(g.B[1], TY);
    const int nbt = XD == 2 ? g.nblk / g.nb[0] : g.nblk;  // blocks (sets) per tile grid
    const int zc = gol_chunk(ctx, (long long)ntx * nty * nbt, R);
    const int nzc = ceil_div(g.B[2], zc);
    dim3 grid(ntx * nty * nzc * nbt);
    gol_bits_kernel<R, NW, NR, IO, RT, XD><<<grid, gb_threads(NW, NR, RT, XD), 0, ctx->stream>>>(
        (const uint8_t*)in, (uint8_t*)out, g, ntx, nty, nzc, zc, pull,
        IO ? gol_pgeo(g) : PGeo{});
    B2R_LAUNCH_CHECK();
    return B2R_OK;
}

template <int R, int IO>
static int gol_bits_launch_r(b2r_ctx* ctx, const void* in, void* out)
{
#define GB_LAUNCH(i)                                                                        \
    gol_bits_launch_t<R, GB_SHAPES[i][0], GB_SHAPES[i][1], GB_SHAPES[i][2], GB_SHAPES[i][3], IO>( \
        ctx, in, out)
    switch (ctx->tune.gol_tile) {
    case 4: return GB_LAUNCH(0);
    case 6: return GB_LAUNCH(2);
    case 7: return GB_LAUNCH(3);
    case 8: return GB_LAUNCH(4);
    case 9: return GB_LAUNCH(5);
    case 10: return GB_LAUNCH(6);
    case 11: return GB_LAUNCH(7);
    case 12: return GB_LAUNCH(8);
    case 13: return gol_whole_rows(ctx->k) ? GB_LAUNCH(9) : GB_LAUNCH(7);
    case 14: return gol_whole_rows(ctx->k) ? GB_LAUNCH(10) : GB_LAUNCH(7);
    case 15: return gol_whole_rows(ctx->k) ? GB_LAUNCH(11) : GB_LAUNCH(7);
    default: return GB_LAUNCH(1);
    }
#undef GB_LAUNCH
}

template <int R>
static int gol_fused_launch_r(b2r_ctx* ctx, const void* in, void* out, bool packed)
{
    if (packed) return gol_bits_launch_r<R, 3>(ctx, in, out);
    switch (ctx->tune.gol_tile) {
    case 1: return gol_fused_launch_t<R, GF_SHAPES[0][0], GF_SHAPES[0][1]>(ctx, in, out);
    case 2: return gol_fused_launch_t<R, GF_SHAPES[1][0], GF_SHAPES[1][1]>(ctx, in, out);
    case 3: return gol_fused_launch_t<R, GF_SHAPES[2][0], GF_SHAPES[2][1]>(ctx, in, out);
    default: return gol_bits_launch_r<R, 0>(ctx, in, out);
    }
}

int b2r_gol_fused_packed(const b2r_ctx* ctx)
{
    return b2r_gol_fused_reads_halo(ctx) && !ctx->tune.gol_bytes;
}

// gens (1..4, at most R) generations of the bit-sliced kernel, io as the
// kernel's IO.  A run's last, partial window (gens < R) runs here as well: the
// halo it needs (gens cells deep) is pulled from the neighbours' owned cells
// as in a whole window, and its update boxes, those of the window's first
// stages, are the widest ones.
template <int R>
static int gol_bits_io_launch_r(b2r_ctx* ctx, const void* in, void* out, int io)
{
    switch (io) {
    case 0: return gol_bits_launch_r<R, 0>(ctx, in, out);
    case 1: return gol_bits_launch_r<R, 1>(ctx, in, out);
    case 2: return gol_bits_launch_r<R, 2>(ctx, in, out);
    default: return gol_bits_launch_r<R, 3>(ctx, in, out);
    }
}

static int gol_bits_packed_launch(b2r_ctx* ctx, const void* in, void* out, int gens, int io)
{
    switch (gens) {
    case 1: return gol_bits_io_launch_r<1>(ctx, in, out, io);
    case 2: return gol_bits_io_launch_r<2>(ctx, in, out, io);
    case 3: return gol_bits_io_launch_r<3>(ctx, in, out, io);
    case 4: return gol_bits_io_launch_r<4>(ctx, in, out, io);
    }
    b2r_set_error("bit-sliced Game of Life runs 1..4 generations per launch (got %d)", gens);
    return B2R_E_UNSUPPORTED;
}

int b2r_gol_pack_pair(b2r_ctx* ctx, const void* from, void** p0, void** p1)
{
    B2R_CHECK(b2r_gol_fused_packed(ctx), B2R_E_UNSUPPORTED,
              "packed Game of Life fields need the bit-sliced kernel and B_x %% 32 == 0");
    const KGeo& g = ctx->k;
    const PGeo pg = gol_pgeo(g);
    const size_t half = ((size_t)pg.bb * g.nblk + 255) & ~(size_t)255;
    if (ctx->gol_pk_bytes < 2 * half) {
        if (ctx->gol_pk) {
            if (int rc = b2r_dev_free(ctx, ctx->gol_pk)) return rc;
            ctx->gol_pk = nullptr;
            ctx->gol_pk_bytes = 0;
        }
        if (int rc = b2r_dev_alloc(ctx, 2 * half, &ctx->gol_pk)) return rc;
        ctx->gol_pk_bytes = 2 * half;
    }
    *p0 = ctx->gol_pk;
    *p1 = (uint8_t*)ctx->gol_pk + half;
    if (!from) return B2R_OK;  // the pair only, contents left as they are
    const long long words = (long long)g.nblk * g.S[2] * g.S[1] * pg.pw;
    gol_pack_kernel<<<(unsigned)ceil_div(words, 256), 256, 0, ctx->stream>>>(
        (const uint8_t*)from, (uint32_t*)*p0, (uint32_t*)*p1, g, pg.pw);
    B2R_LAUNCH_CHECK();
    return B2R_OK;
}

int b2r_gol_fused_launch(b2r_ctx* ctx, const void* in, void* out, int packed)
{
    switch (ctx->g.R) {
    case 1: return gol_fused_launch_r<1>(ctx, in, out, packed);
    case 2: return gol_fused_launch_r<2>(ctx, in, out, packed);
    case 3: return gol_fused_launch_r<3>(ctx, in, out, packed);
    case 4: return gol_fused_launch_r<4>(ctx, in, out, packed);
    }
    b2r_set_error("fused Game of Life supports R <= 4 (got %d)", ctx->g.R);
    return B2R_E_UNSUPPORTED;
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
    int zc = ctx->tune.gol_step_zc;
    if (zc <= 0) {
        long long tiles = (long long)ntx * nty * g.nblk;
        int nzc = (int)((4096 + tiles - 1) / tiles);
        zc = ceil_div(nz, nzc < 1 ? 1 : nzc);
        if (zc < 8) zc = 8;
    }
    if (zc > nz) zc = nz;
    const int nzc = ceil_div(nz, zc);
    dim3 grid(ntx * nty * nzc * g.nblk);
    dim3 block(GOL_TXD, GOL_WY);
    gol_step_kernel<RY><<<grid, block, 0, ctx->stream>>>((const uint8_t*)in, (uint8_t*)out, g, i,
                                                         jd0, ntx, nty, nzc, zc);
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
    const bool use_fused = fused && gol_fused_ok(ctx, R);
    int s = 0;
    if (use_fused && nsteps >= 1 && b2r_gol_fused_packed(ctx)) {
        // The windows run on bit-packed copies of the field: the whole
        // windows, then the remaining nsteps mod R generations in one more
        // launch.  The last launch unpacks as it stores, into the byte field
        // that holds the result.  Where no face is frozen the first launch
        // packs its byte input as it loads it (then no packed halo word is
        // ever read: each comes from the neighbour that owns it); otherwise a
        // pack pass fills both packed fields first, the frozen frame included.
        const int nwin = nsteps / R, tail = nsteps - nwin * R;
        const int nlaunch = nwin + (tail > 0);
        const bool pack_in_first = ctx->k.fmask == 0;
        void* const result = (nlaunch & 1) ? d_b : d_a;
        void* pk[2];
        if (int rc = b2r_gol_pack_pair(ctx, pack_in_first ? nullptr : cur, &pk[0], &pk[1]))
            return rc;
        for (int w = 0; w < nlaunch; ++w) {
            const int gens = w < nwin ? R : tail;
            const bool bytes_in = w == 0 && pack_in_first, bytes_out = w == nlaunch - 1;
            const int io = (bytes_in ? 0 : 1) | (bytes_out ? 0 : 2);
            const void* in = bytes_in ? d_a : pk[w & 1];
            void* out = bytes_out ? result : pk[(w + 1) & 1];
            if (int rc = gol_bits_packed_launch(ctx, in, out, gens, io)) return rc;
            void* t = cur;
            cur = nxt;
            nxt = t;
        }
        s = nsteps;
    }
    for (; s < nsteps; s += R) {
        const int n = nsteps - s < R ? nsteps - s : R;
        const bool fuse = use_fused && n == R;
        // a bit-sliced launch reads its halo from the neighbours itself;
        // gol_exchange runs the pack/unpack exchange of boundary_updating
        if (fuse && ctx->tune.gol_exchange) {
            if (int rc = b2r_boundary_updating(ctx, s, cur)) return rc;
        } else if (!(fuse && b2r_gol_fused_reads_halo(ctx))) {
            if (int rc = b2r_direct_exchange(ctx, &cur, 1)) return rc;
        }
        if (fuse) {
            if (int rc = b2r_gol_fused_launch(ctx, cur, nxt, 0)) return rc;
            void* t = cur;
            cur = nxt;
            nxt = t;
            continue;
        }
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
