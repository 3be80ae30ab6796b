// b2r_core.hip -- context, field layout, uploads, seeded generators, the
// device-local halo exchange (boundary_updating) and measurement helpers.
//
// The reference keeps its geometry in globals (b2r.h:72-87), checks it in
// b2r.cpp:154-167 and simu_env_init (b2r.cpp:47-67), allocates message buffers in
// b2r_sync_init (b2r_halo_sync.cu:74-111) and exchanges one field per call with
// pack -> D2H -> MPI -> H2D -> unpack per dimension (b2r_halo_sync.cu:442-508).
// Here every block of a field is resident on one GPU: one launch per dimension
// packs the faces of all blocks into device message buffers and one launch
// unpacks them, or (the direct path) one launch per dimension copies each halo
// cell straight from its source block for several fields at once.
#include "b2r_common.h"

#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

static thread_local char g_err[512] = "";

void b2r_set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char* b2r_last_error(void) { return g_err; }
extern "C" int b2r_version(void) { return 100; }

int b2r_activate(b2r_ctx* ctx)
{
    B2R_HIP(hipSetDevice(ctx->device));
    return B2R_OK;
}

int b2r_check_field(const b2r_ctx* ctx, const void* f)
{
    B2R_CHECK(f != nullptr, B2R_E_ARG, "null field pointer");
    B2R_CHECK(ctx->fields.count(const_cast<void*>(f)) != 0, B2R_E_ARG,
              "pointer %p is not a field of this context", f);
    return B2R_OK;
}

int b2r_time_begin(b2r_ctx* ctx, float* ms)
{
    if (ms) B2R_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    return B2R_OK;
}

int b2r_time_end(b2r_ctx* ctx, float* ms)
{
    if (ms) {
        B2R_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        B2R_HIP(hipEventSynchronize(ctx->ev1));
        B2R_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    }
    return B2R_OK;
}

// ---------------------------------------------------------------------------
// scratch of b2r_boundary_updating_noalloc: one array per GPU, its address
// taken once per device (b2r_create) and kept here; the array's name is a
// host-side stand-in, so host code passes the kept address, never the name
// ---------------------------------------------------------------------------
constexpr size_t B2R_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[B2R_SCRATCH_BYTES];
constexpr int B2R_MAX_DEVICES = 64;
static void* g_scratch_ptr[B2R_MAX_DEVICES];
static std::mutex g_scratch_mu;

// With `device` current: its scratch address, taken on the first call.
static int scratch_setup(int device)
{
    B2R_CHECK(device >= 0 && device < B2R_MAX_DEVICES, B2R_E_ARG, "device %d beyond %d", device,
              B2R_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return B2R_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    B2R_CHECK(e == hipSuccess && ptr, B2R_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s",
              device, hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return B2R_OK;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
static long long round_up(long long a, long long m) { return (a + m - 1) / m * m; }

extern "C" int b2r_create(const b2r_geom* geom, int device, b2r_ctx** out)
{
    B2R_CHECK(geom && out, B2R_E_ARG, "null argument");
    *out = nullptr;
    const b2r_geom& g = *geom;
    int env[3] = {g.env_x, g.env_y, g.env_z};
    int bd[3] = {g.block_dim_x, g.block_dim_y, g.block_dim_z};
    B2R_CHECK(g.elem_size == 1 || g.elem_size == 4 || g.elem_size == 8, B2R_E_ARG,
              "elem_size must be 1, 4 or 8 (got %d)", g.elem_size);
    B2R_CHECK(g.boundary == B2R_BOUNDARY_FROZEN || g.boundary == B2R_BOUNDARY_PERIODIC,
              B2R_E_ARG, "unknown boundary kind %d", g.boundary);
    B2R_CHECK(g.R >= 1, B2R_E_GEOM, "R must be >= 1 (got %d)", g.R);
    B2R_CHECK(g.radius >= 1, B2R_E_GEOM, "radius must be >= 1 (got %d)", g.radius);
    for (int a = 0; a < 3; ++a) {
        B2R_CHECK(env[a] >= 1 && bd[a] >= 1, B2R_E_GEOM, "extent and block counts must be >= 1");
        // b2r.cpp:154-159: the environment must divide evenly into blocks
        B2R_CHECK(env[a] % bd[a] == 0, B2R_E_GEOM,
                  "ENV_DIM (%d) not divisible by BLOCK_DIM (%d) on axis %d", env[a], bd[a], a);
    }
    int B[3] = {env[0] / bd[0], env[1] / bd[1], env[2] / bd[2]};
    int h = g.radius * g.R;
    // env_z = 1: the 2D environment of _ENV_3D_ = 0 (b2r_config.h:13).  The
    // reference then skips the z check and has no z halo (b2r.cpp:53-67,
    // 142-146); here the one plane keeps a z frame of h planes that stays
    // zero (the z faces are frozen, never periodic, and never exchanged), so
    // every z difference of the 3D kernels vanishes at the plane.
    const bool two_d = env[2] == 1;
    B2R_CHECK(!two_d || bd[2] == 1, B2R_E_GEOM, "a 2D grid (env_z = 1) has BLOCK_DIM_Z = 1");
    int bmin = B[0] < B[1] ? B[0] : B[1];
    if (!two_d) bmin = bmin < B[2] ? bmin : B[2];
    // b2r.cpp:54-59: 1 <= R*D <= min(B)/2 (over x and y only in 2D)
    B2R_CHECK(h <= bmin / 2, B2R_E_GEOM,
              "should meet 1 <= R*radius <= min(B)/2 (R*radius=%d, min(B)=%d)", h, bmin);

    int ndev = 0;
    B2R_HIP(hipGetDeviceCount(&ndev));
    B2R_CHECK(device >= 0 && device < ndev, B2R_E_ARG, "device %d of %d", device, ndev);
    B2R_HIP(hipSetDevice(device));
    if (int rc = scratch_setup(device)) return rc;

    b2r_ctx* c = new b2r_ctx();
    c->g = g;
    c->device = device;
    c->msgbuf = nullptr;
    c->msgbuf_bytes = 0;
    c->gol_pk = nullptr;
    c->gol_pk_bytes = 0;
    memset(&c->tune, 0, sizeof(c->tune));
    b2r_layout& L = c->L;
    memset(&L, 0, sizeof(L));
    for (int a = 0; a < 3; ++a) {
        L.B[a] = B[a];
        L.S[a] = B[a] + 2 * h;
    }
    L.halo = h;
    L.nblocks = bd[0] * bd[1] * bd[2];
    L.elem_size = g.elem_size;
    if (g.elem_size == 1) {
        // byte cells: owned x = h lands on a 16-byte boundary, rows padded to 16
        // bytes with 3 spare bytes, so any dword touching [0,S_x) and any
        // 16-cell group on the grid of x = h is in the row.
        L.x_offset = (16 - h % 16) % 16;
        L.pitch = (int)round_up(L.x_offset + L.S[0] + 3, 16);
    } else {
        L.x_offset = 0;
        L.pitch = L.S[0];
    }
    L.slice = (long long)L.pitch * L.S[1];
    L.block_stride = round_up(L.slice * L.S[2], 256 / g.elem_size);
    L.field_elems = L.block_stride * L.nblocks;

    KGeo& k = c->k;
    for (int a = 0; a < 3; ++a) {
        k.B[a] = L.B[a];
        k.S[a] = L.S[a];
        k.nb[a] = bd[a];
    }
    k.h = h;
    k.two_d = two_d ? 1 : 0;
    k.zclip = 0;
    k.periodic = g.boundary == B2R_BOUNDARY_PERIODIC ? (two_d ? 3 : 7) : 0;
    k.fmask = k.periodic ? (two_d ? 0x30 : 0) : 0x3F;
    k.pitch = L.pitch;
    k.xoff = L.x_offset;
    k.slice = L.slice;
    k.bstride = L.block_stride;
    k.nblk = L.nblocks;

    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
        b2r_set_error("stream/event creation failed: %s", hipGetErrorString(e));
        delete c;
        return B2R_E_HIP;
    }
    *out = c;
    return B2R_OK;
}

extern "C" int b2r_destroy(b2r_ctx* ctx)
{
    if (!ctx) return B2R_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (void* p : ctx->fields) (void)hipFree(p);
    for (void* p : ctx->scratch) (void)hipFree(p);
    if (ctx->msgbuf) (void)hipFree(ctx->msgbuf);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return B2R_OK;
}

extern "C" int b2r_get_layout(const b2r_ctx* ctx, b2r_layout* out)
{
    B2R_CHECK(ctx && out, B2R_E_ARG, "null argument");
    *out = ctx->L;
    return B2R_OK;
}

extern "C" int b2r_synchronize(b2r_ctx* ctx)
{
    B2R_CHECK(ctx, B2R_E_ARG, "null context");
    B2R_HIP(hipStreamSynchronize(ctx->stream));
    return B2R_OK;
}

extern "C" int b2r_set_tuning(b2r_ctx* ctx, const b2r_tuning* t)
{
    B2R_CHECK(ctx && t, B2R_E_ARG, "null argument");
    B2R_CHECK(t->heat_ry == 0 || t->heat_ry == 1 || t->heat_ry == 2 || t->heat_ry == 4 ||
              t->heat_ry == 8, B2R_E_ARG, "heat_ry must be 0, 1, 2, 4 or 8");
    B2R_CHECK(t->heat_zc >= 0 && t->gol_zc >= 0 && t->gol_step_zc >= 0 && t->el_zc >= 0 &&
                  t->el_step_zc >= 0, B2R_E_ARG,
              "negative z chunk");
    B2R_CHECK(t->heat_wy == 0 || t->heat_wy == 2 || t->heat_wy == 3 || t->heat_wy == 4 ||
                  t->heat_wy == 8 || t->heat_wy == 16,
              B2R_E_ARG, "heat_wy must be 0, 2, 3, 4, 8 or 16");
    B2R_CHECK(t->gol_tile >= 0 && t->gol_tile <= 15, B2R_E_ARG, "gol_tile must be 0..15");
    B2R_CHECK(t->el_kernel >= 0 && t->el_kernel <= 2, B2R_E_ARG, "el_kernel must be 0, 1 or 2");
    B2R_CHECK(t->el_grad_mb >= 0, B2R_E_ARG, "el_grad_mb must be >= 0");
    B2R_CHECK(t->gol_bytes == 0 || t->gol_bytes == 1, B2R_E_ARG, "gol_bytes must be 0 or 1");
    B2R_CHECK(t->gol_exchange == 0 || t->gol_exchange == 1, B2R_E_ARG,
              "gol_exchange must be 0 or 1");
    {
        // the (heat_wy, heat_wx) shapes the heat step is built for
        const int wy = t->heat_wy > 0 ? t->heat_wy : 4, wx = t->heat_wx > 0 ? t->heat_wx : 1;
        const bool ok = (wy == 4 && (wx == 1 || wx == 2 || wx == 3 || wx == 4)) ||
                        (wy == 8 && (wx == 1 || wx == 2)) || (wy == 16 && wx == 1) ||
                        (wy == 2 && wx == 6) || (wy == 3 && wx == 3);
        B2R_CHECK(ok && t->heat_wx >= 0, B2R_E_ARG,
                  "(heat_wy, heat_wx) must be (4, 1|2|3|4), (8, 1|2), (16, 1), (2, 6) or (3, 3)");
    }
    ctx->tune = *t;
    return B2R_OK;
}

extern "C" int b2r_get_tuning(const b2r_ctx* ctx, b2r_tuning* t)
{
    B2R_CHECK(ctx && t, B2R_E_ARG, "null argument");
    *t = ctx->tune;
    return B2R_OK;
}

extern "C" int b2r_set_frozen_faces(b2r_ctx* ctx, int mask)
{
    B2R_CHECK(ctx, B2R_E_ARG, "null context");
    B2R_CHECK(mask >= 0 && mask <= 0x3F, B2R_E_ARG, "mask 0x%x outside 6 bits", mask);
    for (int a = 0; a < 3; ++a)
        B2R_CHECK(!((ctx->k.periodic >> a) & 1) || !((mask >> (2 * a)) & 3), B2R_E_ARG,
                  "axis %d is periodic: it has no global faces", a);
    // the z frame of a 2D grid stays frozen
    ctx->k.fmask = ctx->k.two_d ? (mask | 0x30) : mask;
    return B2R_OK;
}

extern "C" int b2r_sync_flag(const b2r_ctx* ctx, int rank, signed char flags[6])
{
    B2R_CHECK(ctx && flags, B2R_E_ARG, "null argument");
    B2R_CHECK(rank >= 0 && rank < ctx->k.nblk, B2R_E_ARG, "rank %d out of range", rank);
    int c[3];
    block_coords(ctx->k, rank, c);
    for (int dim = 0; dim < 3; ++dim)
        for (int side = 0; side < 2; ++side) {
            int n = neighbour_block(ctx->k, rank, dim, side);
            bool wraps = side ? c[dim] == ctx->k.nb[dim] - 1 : c[dim] == 0;
            flags[2 * dim + side] = n < 0 ? 0 : (wraps ? -1 : 1);
        }
    return B2R_OK;
}

// ---------------------------------------------------------------------------
// fields
// ---------------------------------------------------------------------------
template <typename T>
__global__ void fill_kernel(T* __restrict__ f, long long n, T v)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        f[i] = v;
}

static int launch_blocks(long long n, int per = 256, int cap = 65536)
{
    long long b = (n + per - 1) / per;
    if (b < 1) b = 1;
    return (int)(b > cap ? cap : b);
}

extern "C" int b2r_field_alloc(b2r_ctx* ctx, void** d_field)
{
    B2R_CHECK(ctx && d_field, B2R_E_ARG, "null argument");
    if (int rc = b2r_activate(ctx)) return rc;
    size_t bytes = (size_t)ctx->L.field_elems * ctx->g.elem_size;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        b2r_set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return B2R_E_NOMEM;
    }
    B2R_HIP(hipMemsetAsync(p, 0, bytes, ctx->stream));
    ctx->fields.insert(p);
    *d_field = p;
    return B2R_OK;
}

extern "C" int b2r_field_free(b2r_ctx* ctx, void* d_field)
{
    B2R_CHECK(ctx, B2R_E_ARG, "null context");
    if (int rc = b2r_check_field(ctx, d_field)) return rc;
    B2R_HIP(hipStreamSynchronize(ctx->stream));
    B2R_HIP(hipFree(d_field));
    ctx->fields.erase(d_field);
    return B2R_OK;
}

extern "C" int b2r_field_fill(b2r_ctx* ctx, void* d_field, const void* value)
{
    B2R_CHECK(ctx && value, B2R_E_ARG, "null argument");
    if (int rc = b2r_check_field(ctx, d_field)) return rc;
    long long n = ctx->L.field_elems;
    int nb = launch_blocks(n);
    switch (ctx->g.elem_size) {
    case 1: fill_kernel<<<nb, 256, 0, ctx->stream>>>((uint8_t*)d_field, n, *(const uint8_t*)value); break;
    case 4: fill_kernel<<<nb, 256, 0, ctx->stream>>>((uint32_t*)d_field, n, *(const uint32_t*)value); break;
    default: fill_kernel<<<nb, 256, 0, ctx->stream>>>((uint64_t*)d_field, n, *(const uint64_t*)value); break;
    }
    B2R_LAUNCH_CHECK();
    return B2R_OK;
}

// dense global grid <-> owned cells of all blocks
template <typename T, bool TO_FIELD>
__global__ void owned_copy_kernel(T* __restrict__ f, T* __restrict__ dense, KGeo g)
{
    long long gx = (long long)g.B[0] * g.nb[0], gy = (long long)g.B[1] * g.nb[1];
    long long n = gx * gy * (long long)g.B[2] * g.nb[2];
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        int x = (int)(i % gx);
        int y = (int)((i / gx) % gy);
        int z = (int)(i / (gx * gy));
        int b = ((z / g.B[2]) * g.nb[1] + y / g.B[1]) * g.nb[0] + x / g.B[0];
        long long c = cell_index(g, b, g.h + x % g.B[0], g.h + y % g.B[1], g.h + z % g.B[2]);
        if (TO_FIELD) f[c] = dense[i];
        else dense[i] = f[c];
    }
}

template <bool TO_FIELD>
static int owned_copy(b2r_ctx* ctx, void* f, void* dense)
{
    long long n = (long long)ctx->g.env_x * ctx->g.env_y * ctx->g.env_z;
    int nb = launch_blocks(n);
    switch (ctx->g.elem_size) {
    case 1: owned_copy_kernel<uint8_t, TO_FIELD><<<nb, 256, 0, ctx->stream>>>((uint8_t*)f, (uint8_t*)dense, ctx->k); break;
    case 4: owned_copy_kernel<uint32_t, TO_FIELD><<<nb, 256, 0, ctx->stream>>>((uint32_t*)f, (uint32_t*)dense, ctx->k); break;
    default: owned_copy_kernel<uint64_t, TO_FIELD><<<nb, 256, 0, ctx->stream>>>((uint64_t*)f, (uint64_t*)dense, ctx->k); break;
    }
    B2R_LAUNCH_CHECK();
    return B2R_OK;
}

extern "C" int b2r_field_upload(b2r_ctx* ctx, void* d_field, const void* h_global)
{
    B2R_CHECK(ctx && h_global, B2R_E_ARG, "null argument");
    if (int rc = b2r_check_field(ctx, d_field)) return rc;
    if (int rc = b2r_activate(ctx)) return rc;
    size_t bytes = (size_t)ctx->g.env_x * ctx->g.env_y * ctx->g.env_z * ctx->g.elem_size;
    void* tmp = nullptr;
    B2R_HIP(hipMalloc(&tmp, bytes));
    int rc = B2R_OK;
    if (hipMemcpyAsync(tmp, h_global, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
        b2r_set_error("upload copy failed");
        rc = B2R_E_HIP;
    }
    if (!rc) rc = owned_copy<true>(ctx, d_field, tmp);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp);
    return rc;
}

extern "C" int b2r_field_download(b2r_ctx* ctx, const void* d_field, void* h_global)
{
    B2R_CHECK(ctx && h_global, B2R_E_ARG, "null argument");
    if (int rc = b2r_check_field(ctx, d_field)) return rc;
    if (int rc = b2r_activate(ctx)) return rc;
    size_t bytes = (size_t)ctx->g.env_x * ctx->g.env_y * ctx->g.env_z * ctx->g.elem_size;
    void* tmp = nullptr;
    B2R_HIP(hipMalloc(&tmp, bytes));
    int rc = owned_copy<false>(ctx, const_cast<void*>(d_field), tmp);
    if (!rc && hipMemcpyAsync(h_global, tmp, bytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) {
        b2r_set_error("download copy failed");
        rc = B2R_E_HIP;
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp);
    return rc;
}

static int block_copy(b2r_ctx* ctx, void* d_field, int block, void* h_block, bool up)
{
    B2R_CHECK(ctx && h_block, B2R_E_ARG, "null argument");
    if (int rc = b2r_check_field(ctx, d_field)) return rc;
    B2R_CHECK(block >= 0 && block < ctx->k.nblk, B2R_E_ARG, "block %d out of range", block);
    if (int rc = b2r_activate(ctx)) return rc;
    size_t es = ctx->g.elem_size;
    char* dev = (char*)d_field + (size_t)cell_index(ctx->k, block, 0, 0, 0) * es;
    size_t w = (size_t)ctx->L.S[0] * es;
    size_t rows = (size_t)ctx->L.S[1] * ctx->L.S[2];
    if (up)
        B2R_HIP(hipMemcpy2DAsync(dev, (size_t)ctx->L.pitch * es, h_block, w, w, rows,
                                 hipMemcpyHostToDevice, ctx->stream));
    else
        B2R_HIP(hipMemcpy2DAsync(h_block, w, dev, (size_t)ctx->L.pitch * es, w, rows,
                                 hipMemcpyDeviceToHost, ctx->stream));
    B2R_HIP(hipStreamSynchronize(ctx->stream));
    return B2R_OK;
}

extern "C" int b2r_block_upload(b2r_ctx* ctx, void* d_field, int block, const void* h_block)
{
    return block_copy(ctx, d_field, block, const_cast<void*>(h_block), true);
}

extern "C" int b2r_block_download(b2r_ctx* ctx, const void* d_field, int block, void* h_block)
{
    return block_copy(ctx, const_cast<void*>(d_field), block, h_block, false);
}

extern "C" int b2r_field_copy(b2r_ctx* ctx, void* d_dst, const void* d_src)
{
    B2R_CHECK(ctx, B2R_E_ARG, "null context");
    if (int rc = b2r_check_field(ctx, d_dst)) return rc;
    if (int rc = b2r_check_field(ctx, d_src)) return rc;
    B2R_HIP(hipMemcpyAsync(d_dst, d_src, (size_t)ctx->L.field_elems * ctx->g.elem_size,
                           hipMemcpyDeviceToDevice, ctx->stream));
    return B2R_OK;
}

// ---------------------------------------------------------------------------
// seeded generators (owned cells, keyed by the global cell index)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t sm64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// kind 0: uniform f32, kind 1: bernoulli u8
template <int KIND>
__global__ void gen_kernel(void* __restrict__ f, KGeo g, uint64_t key, uint32_t p_q24)
{
    long long gx = (long long)g.B[0] * g.nb[0], gy = (long long)g.B[1] * g.nb[1];
    long long n = gx * gy * (long long)g.B[2] * g.nb[2];
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        int x = (int)(i % gx);
        int y = (int)((i / gx) % gy);
        int z = (int)(i / (gx * gy));
        int b = ((z / g.B[2]) * g.nb[1] + y / g.B[1]) * g.nb[0] + x / g.B[0];
        long long c = cell_index(g, b, g.h + x % g.B[0], g.h + y % g.B[1], g.h + z % g.B[2]);
        uint32_t d = (uint32_t)(sm64(key + (uint64_t)i) >> 40);
        if (KIND == 0) ((float*)f)[c] = (float)d * (1.0f / 16777216.0f);
        else ((uint8_t*)f)[c] = d < p_q24 ? 1 : 0;
    }
}

extern "C" int b2r_field_init_uniform(b2r_ctx* ctx, void* d_field, uint64_t seed)
{
    B2R_CHECK(ctx, B2R_E_ARG, "null context");
    if (int rc = b2r_check_field(ctx, d_field)) return rc;
    B2R_CHECK(ctx->g.elem_size == 4, B2R_E_UNSUPPORTED, "uniform init needs f32 cells");
    uint64_t key = seed;  // host-side splitmix of the seed
    key += 0x9E3779B97F4A7C15ull;
    key = (key ^ (key >> 30)) * 0xBF58476D1CE4E5B9ull;
    key = (key ^ (key >> 27)) * 0x94D049BB133111EBull;
    key = key ^ (key >> 31);
    long long n = (long long)ctx->g.env_x * ctx->g.env_y * ctx->g.env_z;
    gen_kernel<0><<<launch_blocks(n), 256, 0, ctx->stream>>>(d_field, ctx->k, key, 0);
    B2R_LAUNCH_CHECK();
    return B2R_OK;
}

extern "C" int b2r_field_init_bernoulli(b2r_ctx* ctx, void* d_field, uint64_t seed, uint32_t p_q24)
{
    B2R_CHECK(ctx, B2R_E_ARG, "null context");
    if (int rc = b2r_check_field(ctx, d_field)) return rc;
    B2R_CHECK(ctx->g.elem_size == 1, B2R_E_UNSUPPORTED, "bernoulli init needs u8 cells");
    uint64_t key = seed;
    key += 0x9E3779B97F4A7C15ull;
    key = (key ^ (key >> 30)) * 0xBF58476D1CE4E5B9ull;
    key = (key ^ (key >> 27)) * 0x94D049BB133111EBull;
    key = key ^ (key >> 31);
    long long n = (long long)ctx->g.env_x * ctx->g.env_y * ctx->g.env_z;
    gen_kernel<1><<<launch_blocks(n), 256, 0, ctx->stream>>>(d_field, ctx->k, key, p_q24);
    B2R_LAUNCH_CHECK();
    return B2R_OK;
}

__global__ void blobs_kernel(float* __restrict__ f, KGeo g, int nb, const double* __restrict__ amp,
                             const double* __restrict__ ex, const double* __restrict__ ey,
                             const double* __restrict__ ez)
{
    long long gx = (long long)g.B[0] * g.nb[0], gy = (long long)g.B[1] * g.nb[1];
    long long gz = (long long)g.B[2] * g.nb[2];
    long long n = gx * gy * gz;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        int x = (int)(i % gx);
        int y = (int)((i / gx) % gy);
        int z = (int)(i / (gx * gy));
        double acc = 0.0;
        for (int k = 0; k < nb; ++k) {
            double t = __dmul_rn(amp[k], ex[k * gx + x]);
            t = __dmul_rn(t, ey[k * gy + y]);
            t = __dmul_rn(t, ez[k * gz + z]);
            acc = __dadd_rn(acc, t);
        }
        int b = ((z / g.B[2]) * g.nb[1] + y / g.B[1]) * g.nb[0] + x / g.B[0];
        f[cell_index(g, b, g.h + x % g.B[0], g.h + y % g.B[1], g.h + z % g.B[2])] = (float)acc;
    }
}

extern "C" int b2r_field_init_blobs(b2r_ctx* ctx, void* d_field, int nb, const double* amp,
                                    const double* ex, const double* ey, const double* ez)
{
    B2R_CHECK(ctx && amp && ex && ey && ez && nb >= 1, B2R_E_ARG, "bad argument");
    if (int rc = b2r_check_field(ctx, d_field)) return rc;
    B2R_CHECK(ctx->g.elem_size == 4, B2R_E_UNSUPPORTED, "blob init needs f32 cells");
    if (int rc = b2r_activate(ctx)) return rc;
    size_t nx = (size_t)nb * ctx->g.env_x, ny = (size_t)nb * ctx->g.env_y,
           nz = (size_t)nb * ctx->g.env_z;
    double* d = nullptr;
    B2R_HIP(hipMalloc(&d, sizeof(double) * (nb + nx + ny + nz)));
    (void)hipMemcpyAsync(d, amp, sizeof(double) * nb, hipMemcpyHostToDevice, ctx->stream);
    (void)hipMemcpyAsync(d + nb, ex, sizeof(double) * nx, hipMemcpyHostToDevice, ctx->stream);
    (void)hipMemcpyAsync(d + nb + nx, ey, sizeof(double) * ny, hipMemcpyHostToDevice, ctx->stream);
    (void)hipMemcpyAsync(d + nb + nx + ny, ez, sizeof(double) * nz, hipMemcpyHostToDevice, ctx->stream);
    long long n = (long long)ctx->g.env_x * ctx->g.env_y * ctx->g.env_z;
    blobs_kernel<<<launch_blocks(n), 256, 0, ctx->stream>>>((float*)d_field, ctx->k, nb, d, d + nb,
                                                           d + nb + nx, d + nb + nx + ny);
    hipError_t e = hipGetLastError();
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    B2R_CHECK(e == hipSuccess, B2R_E_HIP, "blobs kernel: %s", hipGetErrorString(e));
    return B2R_OK;
}

// ---------------------------------------------------------------------------
// halo exchange
// ---------------------------------------------------------------------------
// Face slab of dim: n0 x n1 in-face cells times h slabs.
//   X: in-face (y,z) over B_Y x B_Z    Y: (x,z) over S_X x B_Z    Z: (x,y) over S_X x S_Y
__host__ __device__ inline void face_dims(const KGeo& g, int dim, int& n0, int& n1)
{
    if (dim == 0) { n0 = g.B[1]; n1 = g.B[2]; }
    else if (dim == 1) { n0 = g.S[0]; n1 = g.B[2]; }
    else { n0 = g.S[0]; n1 = g.S[1]; }
}

// Local coordinates of face cell (j, q0, q1) at slab coordinate a along dim.
__device__ __forceinline__ void face_cell(const KGeo& g, int dim, int a, int q0, int q1,
                                          int& x, int& y, int& z)
{
    if (dim == 0) { x = a; y = g.h + q0; z = g.h + q1; }
    else if (dim == 1) { x = q0; y = a; z = g.h + q1; }
    else { x = q0; y = q1; z = a; }
}

// Message element of (j, q0, q1) in the layouts of b2r_halo_sync.cu:
// X [j][y][z] (z fastest, :210), Y [j][z][x] (:243), Z [j][y][x] (:275).
__device__ __forceinline__ long long msg_elem(int dim, int j, int q0, int q1, int n0, int n1)
{
    if (dim == 0) return ((long long)j * n0 + q0) * n1 + q1;
    return ((long long)j * n1 + q1) * n0 + q0;
}

struct FieldList {
    void* p[16];
};

// Direct exchange along dim: every halo run is copied from the neighbour
// block with V-byte vectors.  X faces: one run of h cells per (y,z) row of the
// owned y,z range.  Y and Z faces: whole rows (the full pitch, row padding
// included, whose contents are unspecified) of the halo rows (y halo x owned z)
// and halo planes, so they move with 16-byte vectors when the pitch allows.
template <int V> struct VecT;
template <> struct VecT<1> { typedef uint8_t T; };
template <> struct VecT<2> { typedef uint16_t T; };
template <> struct VecT<4> { typedef uint32_t T; };
template <> struct VecT<8> { typedef uint2 T; };
template <> struct VecT<16> { typedef uint4 T; };

template <int V>
__global__ void direct_exchange_kernel(FieldList fl, KGeo g, int dim, int esize, int nvec,
                                       long long rows)
{
    typedef typename VecT<V>::T T;
    char* f = (char*)fl.p[blockIdx.y];
    const int h = g.h;
    const long long total = rows * nvec * 2 * g.nblk;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int v = (int)(t % nvec);
        const long long r = t / nvec;
        const long long row = r % rows;
        const int bs = (int)(r / rows);
        const int side = bs & 1, b = bs >> 1;
        const int src = neighbour_block(g, b, dim, side);
        if (src < 0) continue;
        int xd, xs, yd, ys, zd, zs;
        if (dim == 0) {
            yd = ys = h + (int)(row % g.B[1]);
            zd = zs = h + (int)(row / g.B[1]);
            xd = side ? g.B[0] + h : 0;
            xs = side ? h : g.B[0];
        } else if (dim == 1) {
            const int j = (int)(row % h);
            zd = zs = h + (int)(row / h);
            yd = side ? g.B[1] + h + j : j;
            ys = side ? h + j : g.B[1] + j;
            xd = xs = -g.xoff;
        } else {
            yd = ys = (int)(row % g.S[1]);
            const int j = (int)(row / g.S[1]);
            zd = side ? g.B[2] + h + j : j;
            zs = side ? h + j : g.B[2] + j;
            xd = xs = -g.xoff;
        }
        const long long od = cell_index(g, b, xd, yd, zd) * esize + (long long)v * V;
        const long long os = cell_index(g, src, xs, ys, zs) * esize + (long long)v * V;
        *reinterpret_cast<T*>(f + od) = *reinterpret_cast<const T*>(f + os);
    }
}

// widest power-of-two vector (<= 16 bytes) dividing every byte offset in use
static int vec_bytes(long long a, long long b, long long c, long long d)
{
    int v = 16;
    while (v > 1 && ((a % v) || (b % v) || (c % v) || (d % v))) v >>= 1;
    return v;
}

// pack: message (b, side) <- owned slab of block b
template <typename T>
__global__ void pack_kernel(const T* __restrict__ f, T* __restrict__ msgs, KGeo g, int dim,
                            long long me)
{
    int n0, n1;
    face_dims(g, dim, n0, n1);
    const int h = g.h;
    long long total = me * 2 * g.nblk;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        long long r = t / me;
        long long e = t - r * me;
        int side = (int)(r & 1);
        int b = (int)(r >> 1);
        int j, q0, q1;
        if (dim == 0) {
            q1 = (int)(e % n1);
            long long u = e / n1;
            q0 = (int)(u % n0);
            j = (int)(u / n0);
        } else {
            q0 = (int)(e % n0);
            long long u = e / n0;
            q1 = (int)(u % n1);
            j = (int)(u / n1);
        }
        int a = side ? g.B[dim] + j : h + j;
        int x, y, z;
        face_cell(g, dim, a, q0, q1, x, y, z);
        msgs[t] = f[cell_index(g, b, x, y, z)];
    }
}

// unpack: halo slab of block b (side) <- message (neighbour, other side)
template <typename T>
__global__ void unpack_kernel(T* __restrict__ f, const T* __restrict__ msgs, KGeo g, int dim,
                              long long me)
{
    int n0, n1;
    face_dims(g, dim, n0, n1);
    const int h = g.h;
    long long total = me * 2 * g.nblk;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        long long r = t / me;
        long long e = t - r * me;
        int side = (int)(r & 1);
        int b = (int)(r >> 1);
        int src = neighbour_block(g, b, dim, side);
        if (src < 0) continue;
        int j, q0, q1;
        if (dim == 0) {
            q1 = (int)(e % n1);
            long long u = e / n1;
            q0 = (int)(u % n0);
            j = (int)(u / n0);
        } else {
            q0 = (int)(e % n0);
            long long u = e / n0;
            q1 = (int)(u % n1);
            j = (int)(u / n1);
        }
        int a = side ? g.B[dim] + h + j : j;
        int x, y, z;
        face_cell(g, dim, a, q0, q1, x, y, z);
        f[cell_index(g, b, x, y, z)] = msgs[((long long)src * 2 + (1 - side)) * me + e];
    }
}

static long long msg_elems_k(const KGeo& g, int dim)
{
    int n0, n1;
    face_dims(g, dim, n0, n1);
    return (long long)g.h * n0 * n1;
}

extern "C" long long b2r_msg_elems(const b2r_ctx* ctx, int dim)
{
    if (!ctx || dim < 0 || dim > 2) {
        b2r_set_error("bad argument");
        return B2R_E_ARG;
    }
    return msg_elems_k(ctx->k, dim);
}

extern "C" int b2r_pack(b2r_ctx* ctx, int dim, const void* d_field, void* d_msgs)
{
    B2R_CHECK(ctx && d_msgs && dim >= 0 && dim <= 2, B2R_E_ARG, "bad argument");
    if (int rc = b2r_check_field(ctx, d_field)) return rc;
    long long me = msg_elems_k(ctx->k, dim);
    int nb = launch_blocks(me * 2 * ctx->k.nblk);
    switch (ctx->g.elem_size) {
    case 1: pack_kernel<uint8_t><<<nb, 256, 0, ctx->stream>>>((const uint8_t*)d_field, (uint8_t*)d_msgs, ctx->k, dim, me); break;
    case 4: pack_kernel<uint32_t><<<nb, 256, 0, ctx->stream>>>((const uint32_t*)d_field, (uint32_t*)d_msgs, ctx->k, dim, me); break;
    default: pack_kernel<uint64_t><<<nb, 256, 0, ctx->stream>>>((const uint64_t*)d_field, (uint64_t*)d_msgs, ctx->k, dim, me); break;
    }
    B2R_LAUNCH_CHECK();
    return B2R_OK;
}

extern "C" int b2r_unpack(b2r_ctx* ctx, int dim, void* d_field, const void* d_msgs)
{
    B2R_CHECK(ctx && d_msgs && dim >= 0 && dim <= 2, B2R_E_ARG, "bad argument");
    if (int rc = b2r_check_field(ctx, d_field)) return rc;
    long long me = msg_elems_k(ctx->k, dim);
    int nb = launch_blocks(me * 2 * ctx->k.nblk);
    switch (ctx->g.elem_size) {
    case 1: unpack_kernel<uint8_t><<<nb, 256, 0, ctx->stream>>>((uint8_t*)d_field, (const uint8_t*)d_msgs, ctx->k, dim, me); break;
    case 4: unpack_kernel<uint32_t><<<nb, 256, 0, ctx->stream>>>((uint32_t*)d_field, (const uint32_t*)d_msgs, ctx->k, dim, me); break;
    default: unpack_kernel<uint64_t><<<nb, 256, 0, ctx->stream>>>((uint64_t*)d_field, (const uint64_t*)d_msgs, ctx->k, dim, me); break;
    }
    B2R_LAUNCH_CHECK();
    return B2R_OK;
}

extern "C" int b2r_boundary_updating(b2r_ctx* ctx, int time_step, void* d_field)
{
    (void)time_step;
    B2R_CHECK(ctx, B2R_E_ARG, "null context");
    if (int rc = b2r_check_field(ctx, d_field)) return rc;
    if (int rc = b2r_activate(ctx)) return rc;
    // one message buffer sized for the largest direction (b2r_halo_sync.cu:82-84)
    size_t need = 0;
    for (int d = 0; d < 3; ++d) {
        size_t s = (size_t)msg_elems_k(ctx->k, d) * 2 * ctx->k.nblk * ctx->g.elem_size;
        need = s > need ? s : need;
    }
    if (ctx->msgbuf_bytes < need) {
        if (ctx->msgbuf) {
            B2R_HIP(hipStreamSynchronize(ctx->stream));
            B2R_HIP(hipFree(ctx->msgbuf));
            ctx->msgbuf = nullptr;
        }
        B2R_HIP(hipMalloc(&ctx->msgbuf, need));
        ctx->msgbuf_bytes = need;
    }
    // a 2D grid has no Z phase (b2r_halo_sync.cu:170, 489)
    for (int d = 0; d < (ctx->k.two_d ? 2 : 3); ++d) {
        if (int rc = b2r_pack(ctx, d, d_field, ctx->msgbuf)) return rc;
        if (int rc = b2r_unpack(ctx, d, d_field, ctx->msgbuf)) return rc;
    }
    return B2R_OK;
}

// Bytes of the message buffer of one exchange direction (all blocks, both
// sides), the largest over the directions the exchange runs.
static size_t msg_bytes_max(const b2r_ctx* ctx)
{
    size_t need = 0;
    for (int d = 0; d < (ctx->k.two_d ? 2 : 3); ++d) {
        size_t s = (size_t)msg_elems_k(ctx->k, d) * 2 * ctx->k.nblk * ctx->g.elem_size;
        need = s > need ? s : need;
    }
    return need;
}

extern "C" int b2r_boundary_updating_noalloc(b2r_ctx* ctx, int time_step, void* d_field)
{
    (void)time_step;
    B2R_CHECK(ctx, B2R_E_ARG, "null context");
    if (int rc = b2r_check_field(ctx, d_field)) return rc;
    const size_t need = msg_bytes_max(ctx);
    B2R_CHECK(need <= B2R_SCRATCH_BYTES, B2R_E_ARG,
              "boundary_updating_noalloc: the messages of one direction need %zu bytes, more "
              "than the %zu-byte scratch (use b2r_boundary_updating)",
              need, B2R_SCRATCH_BYTES);
    if (int rc = b2r_activate(ctx)) return rc;
    void* msgs = g_scratch_ptr[ctx->device];
    B2R_CHECK(msgs, B2R_E_HIP, "no scratch address for device %d", ctx->device);
    // every message element the unpack reads, the pack of the same direction
    // has just written: nothing needs clearing
    for (int d = 0; d < (ctx->k.two_d ? 2 : 3); ++d) {
        if (int rc = b2r_pack(ctx, d, d_field, msgs)) return rc;
        if (int rc = b2r_unpack(ctx, d, d_field, msgs)) return rc;
    }
    return B2R_OK;
}

int b2r_direct_exchange(b2r_ctx* ctx, void* const* fields, int nfields)
{
    const KGeo& g = ctx->k;
    const long long es = ctx->g.elem_size;
    const long long pb = g.pitch * es, sb = g.slice * es, bb = g.bstride * es;
    for (int base = 0; base < nfields; base += 16) {
        int n = nfields - base < 16 ? nfields - base : 16;
        FieldList fl;
        for (int i = 0; i < n; ++i) fl.p[i] = fields[base + i];
        for (int d = 0; d < (g.two_d ? 2 : 3); ++d) {
            // one block along d and not periodic: no block has a neighbour
            // there, and every thread of the launch would skip its copy
            if (g.nb[d] == 1 && !((g.periodic >> d) & 1)) continue;
            int V;
            long long rows;
            if (d == 0) {
                V = vec_bytes(g.h * es, g.B[0] * es, g.xoff * es, pb);
                V = vec_bytes(V, sb, bb, V);
                rows = (long long)g.B[1] * g.B[2];
            } else {
                V = vec_bytes(pb, sb, bb, 16);
                rows = d == 1 ? (long long)g.h * g.B[2] : (long long)g.h * g.S[1];
            }
            const int nvec = (int)((d == 0 ? g.h * es : pb) / V);
            const long long total = rows * nvec * 2 * g.nblk;
            dim3 grid(launch_blocks(total, 256, 16384), n);
            const int e = (int)es;
            switch (V) {
            case 16: direct_exchange_kernel<16><<<grid, 256, 0, ctx->stream>>>(fl, g, d, e, nvec, rows); break;
            case 8: direct_exchange_kernel<8><<<grid, 256, 0, ctx->stream>>>(fl, g, d, e, nvec, rows); break;
            case 4: direct_exchange_kernel<4><<<grid, 256, 0, ctx->stream>>>(fl, g, d, e, nvec, rows); break;
            case 2: direct_exchange_kernel<2><<<grid, 256, 0, ctx->stream>>>(fl, g, d, e, nvec, rows); break;
            default: direct_exchange_kernel<1><<<grid, 256, 0, ctx->stream>>>(fl, g, d, e, nvec, rows); break;
            }
            B2R_LAUNCH_CHECK();
        }
    }
    return B2R_OK;
}

extern "C" int b2r_boundary_updating_direct(b2r_ctx* ctx, int time_step, void* const* d_fields,
                                            int nfields)
{
    (void)time_step;
    B2R_CHECK(ctx && d_fields && nfields >= 1, B2R_E_ARG, "bad argument");
    for (int i = 0; i < nfields; ++i)
        if (int rc = b2r_check_field(ctx, d_fields[i])) return rc;
    if (int rc = b2r_activate(ctx)) return rc;
    return b2r_direct_exchange(ctx, d_fields, nfields);
}

// ---------------------------------------------------------------------------
// measurement helpers
// ---------------------------------------------------------------------------
// One 16-byte element per thread with non-temporal loads and stores: of the
// copy forms measured on MI355X (grid-stride loops, 4 elements per thread,
// plain or non-temporal single elements) this one moves the most bytes.
__global__ __launch_bounds__(256) void stream_copy_kernel(const uint4* __restrict__ src,
                                                          uint4* __restrict__ dst, size_t n)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 v;
    v.x = __builtin_nontemporal_load(&src[i].x);
    v.y = __builtin_nontemporal_load(&src[i].y);
    v.z = __builtin_nontemporal_load(&src[i].z);
    v.w = __builtin_nontemporal_load(&src[i].w);
    __builtin_nontemporal_store(v.x, &dst[i].x);
    __builtin_nontemporal_store(v.y, &dst[i].y);
    __builtin_nontemporal_store(v.z, &dst[i].z);
    __builtin_nontemporal_store(v.w, &dst[i].w);
}

extern "C" int b2r_stream_copy(b2r_ctx* ctx, void* d_dst, const void* d_src, size_t bytes,
                               int iters, float* ms)
{
    B2R_CHECK(ctx && d_dst && d_src && iters >= 1, B2R_E_ARG, "bad argument");
    B2R_CHECK(bytes % 16 == 0, B2R_E_ARG, "bytes must be a multiple of 16");
    if (int rc = b2r_activate(ctx)) return rc;
    size_t n = bytes / 16;
    B2R_CHECK((n + 255) / 256 <= 0x7fffffffull, B2R_E_ARG, "copy too large");
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (int rc = b2r_time_begin(ctx, ms)) return rc;
    for (int it = 0; it < iters; ++it)
        stream_copy_kernel<<<nb, 256, 0, ctx->stream>>>((const uint4*)d_src, (uint4*)d_dst, n);
    B2R_LAUNCH_CHECK();
    return b2r_time_end(ctx, ms);
}

extern "C" int b2r_dev_alloc(b2r_ctx* ctx, size_t bytes, void** d_ptr)
{
    B2R_CHECK(ctx && d_ptr, B2R_E_ARG, "null argument");
    if (int rc = b2r_activate(ctx)) return rc;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        b2r_set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return B2R_E_NOMEM;
    }
    ctx->scratch.insert(p);
    *d_ptr = p;
    return B2R_OK;
}

extern "C" int b2r_dev_free(b2r_ctx* ctx, void* d_ptr)
{
    B2R_CHECK(ctx && ctx->scratch.count(d_ptr), B2R_E_ARG, "not a scratch allocation");
    B2R_HIP(hipStreamSynchronize(ctx->stream));
    B2R_HIP(hipFree(d_ptr));
    ctx->scratch.erase(d_ptr);
    return B2R_OK;
}

extern "C" int b2r_memcpy_h2d(b2r_ctx* ctx, void* d_dst, const void* h_src, size_t bytes)
{
    B2R_CHECK(ctx && d_dst && h_src, B2R_E_ARG, "null argument");
    B2R_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    B2R_HIP(hipStreamSynchronize(ctx->stream));
    return B2R_OK;
}

extern "C" int b2r_memcpy_d2h(b2r_ctx* ctx, void* h_dst, const void* d_src, size_t bytes)
{
    B2R_CHECK(ctx && h_dst && d_src, B2R_E_ARG, "null argument");
    B2R_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    B2R_HIP(hipStreamSynchronize(ctx->stream));
    return B2R_OK;
}

// Back-to-back launches of one kernel between two events on the context
// stream, alternating the two buffers.  Inner-step kernels cycle b2r_i over
// 0..R-1 so the average covers a whole B2R period; fused kernels run R steps
// per launch.  Fused Game of Life launches that b2r_gol_run makes on packed
// fields alternate the context's packed pair, packed from d_a before the
// timed region.
extern "C" int b2r_time_steps(b2r_ctx* ctx, int kind, void* d_a, void* d_b,
                              const b2r_heat_params* p, int iters, float* ms)
{
    B2R_CHECK(ctx && ms && iters >= 1 && kind >= 0 && kind <= 3, B2R_E_ARG, "bad argument");
    if (int rc = b2r_check_field(ctx, d_a)) return rc;
    if (int rc = b2r_check_field(ctx, d_b)) return rc;
    B2R_CHECK((kind != 0 && kind != 3) || p, B2R_E_ARG, "heat timing needs parameters");
    if (int rc = b2r_activate(ctx)) return rc;
    const int packed = kind == 2 && b2r_gol_fused_packed(ctx);
    if (packed)
        if (int rc = b2r_gol_pack_pair(ctx, d_a, &d_a, &d_b)) return rc;
    if (int rc = b2r_time_begin(ctx, ms)) return rc;
    for (int it = 0; it < iters; ++it) {
        const void* in = (it & 1) ? d_b : d_a;
        void* out = (it & 1) ? d_a : d_b;
        const int i = it % ctx->g.R;
        int rc;
        switch (kind) {
        case 0: rc = b2r_heat_launch(ctx, in, out, p, i); break;
        case 1: rc = b2r_gol_launch(ctx, in, out, i); break;
        case 2: rc = b2r_gol_fused_launch(ctx, in, out, packed); break;
        default: rc = b2r_heat_fused_launch(ctx, in, out, p); break;
        }
        if (rc) return rc;
    }
    return b2r_time_end(ctx, ms);
}
