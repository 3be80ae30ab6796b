// b2r_common.h -- internal definitions shared by the gfx950 translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <set>
#include <string>

#include "../../include/b2r_hip.h"

// ---------------------------------------------------------------------------
// errors: every ABI call returns a code and leaves a message (never exits)
// ---------------------------------------------------------------------------
void b2r_set_error(const char* fmt, ...);

#define B2R_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            b2r_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),  \
                          __FILE__, __LINE__);                                     \
            return B2R_E_HIP;                                                      \
        }                                                                          \
    } while (0)

#define B2R_CHECK(cond, code, ...)                                                 \
    do {                                                                           \
        if (!(cond)) {                                                             \
            b2r_set_error(__VA_ARGS__);                                            \
            return (code);                                                         \
        }                                                                          \
    } while (0)

#define B2R_LAUNCH_CHECK()                                                         \
    do {                                                                           \
        hipError_t e_ = hipGetLastError();                                         \
        if (e_ != hipSuccess) {                                                    \
            b2r_set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_),\
                          __FILE__, __LINE__);                                     \
            return B2R_E_HIP;                                                      \
        }                                                                          \
    } while (0)

// ---------------------------------------------------------------------------
// geometry passed to kernels by value
// ---------------------------------------------------------------------------
struct KGeo {
    int B[3];           // owned cells per block
    int S[3];           // padded block size
    int nb[3];          // blocks per axis
    int h;              // halo depth
    int periodic;       // bit a: the global faces of axis a wrap (x 1, y 2, z 4)
    int fmask;          // bit 2*axis+side: that global face is frozen (clip to owned)
    int two_d;          // 1: one z plane (env_z = 1, _ENV_3D_ = 0 of b2r_config.h:13),
                        // its z halo a frozen zero frame that no exchange touches
    int zclip;          // > 0: update boxes end below global plane zclip - r*(i+1)
                        // (the GPU part of a hybrid run, whose z halo ends there)
    int pitch;          // elements per row
    int xoff;           // elements before x = 0 in a row
    long long slice;    // elements per plane
    long long bstride;  // elements per block
    int nblk;
};

struct b2r_ctx {
    b2r_geom g;
    b2r_layout L;
    KGeo k;
    int device;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    std::set<void*> fields;   // allocations made by b2r_field_alloc
    std::set<void*> scratch;  // allocations made by b2r_dev_alloc
    void* msgbuf;             // boundary_updating message buffers
    size_t msgbuf_bytes;
    b2r_tuning tune;          // launch shapes (0 = automatic)
    void* gol_pk;             // two bit-packed Game of Life fields (b2r_dev_alloc)
    size_t gol_pk_bytes;
};

int b2r_activate(b2r_ctx* ctx);  // hipSetDevice(ctx->device)
int b2r_check_field(const b2r_ctx* ctx, const void* f);
// Ordered X, Y, Z halo exchange inside each field.
int b2r_direct_exchange(b2r_ctx* ctx, void* const* fields, int nfields);
int b2r_time_begin(b2r_ctx* ctx, float* ms);
int b2r_time_end(b2r_ctx* ctx, float* ms);
int b2r_heat_check(b2r_ctx* ctx, const b2r_heat_params* p);
int b2r_heat_launch(b2r_ctx* ctx, const void* in, void* out, const b2r_heat_params* p, int i);
int b2r_gol_launch(b2r_ctx* ctx, const void* in, void* out, int i);
// packed: in and out are bit-packed fields (b2r_gol_pack_pair)
int b2r_gol_fused_launch(b2r_ctx* ctx, const void* in, void* out, int packed);
// 1: a fused Game of Life launch with the context's tuning reads its halo
// from the neighbouring blocks itself (no exchange is needed before it)
int b2r_gol_fused_reads_halo(const b2r_ctx* ctx);
// 1: fused Game of Life runs keep their fields bit-packed between launches
int b2r_gol_fused_packed(const b2r_ctx* ctx);
// 1: gol_tile `tile` runs as itself on this context's geometry (the
// whole-row tiles 13-15 need a grid periodic in x over 512 cells)
int b2r_gol_tile_ok(const b2r_ctx* ctx, int tile);
// the z chunk a fused Game of Life launch takes with the context's tuning
// (gol_zc, or the automatic chunk where that is 0)
int b2r_gol_fused_chunk(const b2r_ctx* ctx);
// The context's two packed fields (allocated on first use), both packed from
// the byte field `from`, halo included (from = NULL: allocated only, contents
// unspecified).
int b2r_gol_pack_pair(b2r_ctx* ctx, const void* from, void** p0, void** p1);
int b2r_heat_fused_launch(b2r_ctx* ctx, const void* in, void* out, const b2r_heat_params* p);

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
// Inner-step update region of block b for a per-step radius r at inner step i:
// [r*(i+1), S-1-r*(i+1)] per axis (heat-3D.cu:191, 268), clipped to the owned
// range at frozen global faces (heat-3D.cu:81-114).  An open global face (its
// fmask bit clear) is updated like an interior one, as the reference's general
// kernels do (heat-3D.cu:173-235), and no exchange ever refreshes it.
struct Box {
    int lo[3], hi[3];
};

__host__ __device__ inline void block_coords(const KGeo& g, int b, int c[3])
{
    c[0] = b % g.nb[0];
    c[1] = (b / g.nb[0]) % g.nb[1];
    c[2] = b / (g.nb[0] * g.nb[1]);
}

__host__ __device__ inline Box update_box(const KGeo& g, int b, int r, int i)
{
    int c[3];
    block_coords(g, b, c);
    Box bx;
    for (int a = 0; a < 3; ++a) {
        const bool per = (g.periodic >> a) & 1;
        bool lo_edge = !per && c[a] == 0 && ((g.fmask >> (2 * a)) & 1);
        bool hi_edge = !per && c[a] == g.nb[a] - 1 && ((g.fmask >> (2 * a + 1)) & 1);
        int lo = r * (i + 1), hi = g.S[a] - 1 - r * (i + 1);
        bx.lo[a] = (lo_edge && g.h > lo) ? g.h : lo;
        bx.hi[a] = (hi_edge && g.h + g.B[a] - 1 < hi) ? g.h + g.B[a] - 1 : hi;
    }
    if (g.zclip > 0) {
        // global plane zclip - 1 - r*(i+1) in this block's local coordinates
        const int hz = g.zclip - 1 - r * (i + 1) - c[2] * g.B[2] + g.h;
        if (hz < bx.hi[2]) bx.hi[2] = hz;
    }
    return bx;
}

static inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// Every global face is either periodic or frozen (none open): the condition of
// the fused kernels, whose redundant halo stages need halos that the exchange
// refreshes or that never change.
__host__ __device__ inline bool closed_faces(const KGeo& g)
{
    for (int a = 0; a < 3; ++a)
        if (!((g.periodic >> a) & 1) && ((g.fmask >> (2 * a)) & 3) != 3) return false;
    return true;
}

// Neighbour block of b along dim on side 0 (minus) or 1 (plus), -1 if the
// block sits on a frozen global face (b2r_sync_flag, b2r_halo_sync.cu:140-183;
// rank map b2r_comm.hpp:25-28; periodic wrap added).
__host__ __device__ inline int neighbour_block(const KGeo& g, int b, int dim, int side)
{
    int c[3];
    block_coords(g, b, c);
    c[dim] += side ? 1 : -1;
    if (c[dim] < 0 || c[dim] >= g.nb[dim]) {
        if (!((g.periodic >> dim) & 1)) return -1;
        c[dim] = (c[dim] + g.nb[dim]) % g.nb[dim];
    }
    return (c[2] * g.nb[1] + c[1]) * g.nb[0] + c[0];
}

// XCD-aware workgroup order.  The dispatcher deals workgroups round-robin over
// the 8 XCDs of an MI355X (linear id mod 8), each with its own L2.  Mapping the
// linear id L (of N) to t so that XCD k receives a contiguous range of t keeps
// neighbouring tiles, which share halo rows and planes, in the same L2.
// A bijection for any N.
__device__ __forceinline__ int xcd_tile(int L, int N)
{
    const int q = N >> 3, r = N & 7;
    const int k = L & 7, i = L >> 3;
    return k * q + (k < r ? k : r) + i;
}

// Global-memory access by 32-bit byte offsets.  A stencil kernel takes a base
// pointer that is uniform over its workgroup (scalar registers) and addresses
// cells by unsigned byte offsets from it, so loads use the scalar-base form
// and one offset register serves every field with the same layout.  The
// caller keeps every offset a kernel forms below 4 GiB.  Casting to the
// global address space also replaces flat instructions where the compiler
// cannot tell that a pointer read from memory is device global memory.
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) char gbyte;
__device__ __forceinline__ gfloat* G(float* p) { return (gfloat*)p; }
__device__ __forceinline__ const gfloat* G(const float* p) { return (const gfloat*)p; }
__device__ __forceinline__ float ldo(const gfloat* __restrict__ p, uint32_t o)
{
    return *(const gfloat*)((const gbyte*)p + o);
}
__device__ __forceinline__ void sto(gfloat* __restrict__ p, uint32_t o, float v)
{
    *(gfloat*)((gbyte*)p + o) = v;
}

// Raw buffer access for marches: a resource over a uniform base pointer (no
// range check: the caller keeps every offset inside the allocation), the
// per-thread byte offset in a vector register and the uniform part (a plane
// offset) in the instruction's scalar offset, so advancing a plane costs no
// vector instruction.  0x00020000 is the gfx9 buffer resource word 3 for
// 32-bit raw accesses.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, -1, 0x00020000);
}
__device__ __forceinline__ float bld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff,
                                    float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
}

// Element offset of local cell (x,y,z) of block b.
__host__ __device__ inline long long cell_index(const KGeo& g, int b, int x, int y, int z)
{
    return (long long)b * g.bstride + g.xoff + (long long)z * g.slice + (long long)y * g.pitch + x;
}
