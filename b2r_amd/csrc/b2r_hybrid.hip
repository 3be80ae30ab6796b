// b2r_hybrid.hip -- the heat step on the host cores, and hybrid runs in which
// the host cores and the GPU advance one field together.
//
// Reference: the CPU steps cpu_fdm_heat_diffuse1..4 (heat-3D.cpp:48-134) and
// their driver heat_3D_cpu_main (:136-184); the "joint layer" that the CPU
// loop's reversed z order protects (heat-3D.cpp:47, 55), and
// heat_3D_gpu_diffsync_main, declared for the CPU+GPU scheme of the B2R paper
// (heat-3D.hpp:15) but never defined.
//
// CPU step: the device's arithmetic restated on the host, term for term
// (b2r_heat_ops.h: F32 mode by fused multiply-adds, REF mode with the FP64
// promotion of heat-3D.cpp:57-60), compiled like the device code without FP
// contraction, so a host step equals a device step bit for bit.  It updates the
// region of the reference's plain step, [D(i+1), N-1-D(i+1)] on every axis
// (the reference's CPU delta2-4 use the bounds and taps of delta1, which is not
// reproduced).
//
// Hybrid run: the host cores own the top cpu_planes owned z planes of the
// global grid (a slab over the whole x-y extent), the GPU the planes below,
// each with a joint layer of h = R*D planes of the other's cells.  Every R
// steps the two exchange their joint layers over PCIe (the GPU's top h owned
// planes down, the slab's bottom h owned planes up), the GPU fields run the
// device exchange between blocks, then both sides take R inner steps at once:
// the GPU's update boxes end h - D(i+1) planes into the slab (KGeo.zclip), the
// slab's region starts D(i+1) planes into its joint layer, as between any two
// B2R blocks.  Every owned cell then sees exactly the inputs of the GPU-only
// run, so the result equals it bit for bit; the slab's planes are written back
// into the result field at the end.
#include "b2r_common.h"
#include "b2r_heat_ops.h"

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <unistd.h>
#include <vector>

// Host worker threads, started on first use and kept until the process ends,
// so a hybrid window pays no thread start-up: run(n, job) calls job(t, n) on
// workers t = 0..n-1 and returns when all have finished.  Calls are serialised.
// The pool is never destroyed (its idle workers end with the process), and a
// process forked after using it starts a pool of its own (host_pool), since
// the parent's workers do not exist in the child.
class HostPool {
public:
    void run(int n, const std::function<void(int, int)>& job)
    {
        std::lock_guard<std::mutex> one(call_);
        {
            std::lock_guard<std::mutex> lk(m_);
            while ((int)th_.size() < n) th_.emplace_back(&HostPool::worker, this, (int)th_.size());
            job_ = &job;
            n_ = n;
            pending_ = (int)th_.size();
            ++gen_;
        }
        go_.notify_all();
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return pending_ == 0; });
    }

private:
    void worker(int id)
    {
        long long seen = 0;
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            go_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            const int n = n_;
            const std::function<void(int, int)>* job = job_;
            lk.unlock();
            if (id < n) (*job)(id, n);
            lk.lock();
            if (--pending_ == 0) done_.notify_one();
        }
    }
    std::mutex call_, m_;
    std::condition_variable go_, done_;
    std::vector<std::thread> th_;
    const std::function<void(int, int)>* job_ = nullptr;
    int n_ = 0, pending_ = 0;
    long long gen_ = 0;
};

static HostPool& host_pool()
{
    static std::mutex mu;
    static HostPool* pool = nullptr;
    static pid_t owner = 0;
    std::lock_guard<std::mutex> lk(mu);
    if (!pool || owner != getpid()) {
        pool = new HostPool();  // a forked child leaves the parent's pool unused
        owner = getpid();
    }
    return *pool;
}

// ---------------------------------------------------------------- host step
template <int D, int MODE>
static inline float cpu_update(const float* c, long long sy, long long sz, const HeatK& hk)
{
    const long long s[3] = {1, sy, sz};
    float L[3];
    const float uc = c[0];
    for (int a = 0; a < 3; ++a) {
        const long long st = s[a];
        if (MODE == B2R_HEAT_REF && D == 1) {
            L[a] = (c[-st] - 2.0f * uc) + c[st];
        } else if (MODE == B2R_HEAT_REF) {
            float acc = heat_c(D, D) * (c[-D * st] + c[D * st]);
            for (int j = D - 1; j >= 1; --j) acc = acc + heat_c(D, j) * (c[-j * st] + c[j * st]);
            L[a] = acc + heat_c(D, 0) * uc;
        } else {
            float acc = heat_c(D, D) * (c[-D * st] + c[D * st]);
            for (int j = D - 1; j >= 1; --j) acc = std::fmaf(heat_c(D, j), c[-j * st] + c[j * st], acc);
            L[a] = std::fmaf(heat_c(D, 0), uc, acc);
        }
    }
    if (MODE == B2R_HEAT_F32) {
        float t = hk.k[0] * L[0];
        t = std::fmaf(hk.k[1], L[1], t);
        t = std::fmaf(hk.k[2], L[2], t);
        return uc + t;
    }
    double t = (double)L[0] / hk.h2[0];
    t = t + (double)L[1] / hk.h2[1];
    t = t + (double)L[2] / hk.h2[2];
    return (float)((double)uc + hk.adt * t);
}

// out = step(in) on the box lo..hi (inclusive, per axis) of a dense
// Nx*Ny*Nz grid, x fastest; the box's (z, y) rows shared out over the host
// pool's workers in contiguous ranges, so a slab of a few planes still uses
// every worker.
template <int D, int MODE>
static void cpu_box(const float* in, float* out, int Nx, int Ny, const int lo[3], const int hi[3],
                    const HeatK& hk, int nthreads)
{
    const long long sy = Nx, sz = (long long)Nx * Ny;
    const int nz = hi[2] - lo[2] + 1, ny = hi[1] - lo[1] + 1;
    if (nz <= 0 || ny <= 0 || hi[0] < lo[0]) return;
    const long long rows = (long long)nz * ny;
    // the reference walks z downwards (heat-3D.cpp:55); the result does not
    // depend on the order (the step reads in and writes out)
    auto work = [&](int t, int nt) {
        const long long r0 = rows * t / nt, r1 = rows * (t + 1) / nt;
        for (long long r = r0; r < r1; ++r) {
            const int z = hi[2] - (int)(r / ny), y = lo[1] + (int)(r % ny);
            const long long o = z * sz + y * sy;
            for (int x = lo[0]; x <= hi[0]; ++x)
                out[o + x] = cpu_update<D, MODE>(in + o + x, sy, sz, hk);
        }
    };
    int nt = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > rows) nt = (int)rows;
    if (nt == 1) {
        work(0, 1);
        return;
    }
    host_pool().run(nt, work);
}

static void cpu_box_any(const b2r_heat_params* p, const float* in, float* out, int Nx, int Ny,
                        const int lo[3], const int hi[3], int nthreads)
{
    const HeatK hk = heat_k(p);
#define B2R_CPU_CASE(DD)                                                                     \
    case DD:                                                                                 \
        if (p->mode == B2R_HEAT_F32) cpu_box<DD, B2R_HEAT_F32>(in, out, Nx, Ny, lo, hi, hk, nthreads); \
        else cpu_box<DD, B2R_HEAT_REF>(in, out, Nx, Ny, lo, hi, hk, nthreads);            \
        break;
    switch (p->D) {
        B2R_CPU_CASE(1)
        B2R_CPU_CASE(2)
        B2R_CPU_CASE(3)
        default: B2R_CPU_CASE(4)
    }
#undef B2R_CPU_CASE
}

extern "C" int b2r_heat_cpu_step(const b2r_heat_params* p, const float* h_in, float* h_out,
                                 int Ngx, int Ngy, int Ngz, int b2r_i, int nthreads)
{
    B2R_CHECK(p && h_in && h_out && h_in != h_out, B2R_E_ARG, "bad argument");
    B2R_CHECK(p->D >= 1 && p->D <= 4, B2R_E_UNSUPPORTED, "unsupported delta D=%d", p->D);
    B2R_CHECK(p->mode == B2R_HEAT_F32 || p->mode == B2R_HEAT_REF, B2R_E_ARG, "bad mode");
    B2R_CHECK(Ngx >= 1 && Ngy >= 1 && Ngz >= 1 && b2r_i >= 0, B2R_E_ARG, "bad grid or step");
    if (p->mode == B2R_HEAT_REF)
        B2R_CHECK(p->h2[0] > 0 && p->h2[1] > 0 && p->h2[2] > 0, B2R_E_ARG, "h2 must be > 0");
    const int N[3] = {Ngx, Ngy, Ngz};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = p->D * (b2r_i + 1);
        hi[a] = N[a] - 1 - p->D * (b2r_i + 1);
    }
    cpu_box_any(p, h_in, h_out, Ngx, Ngy, lo, hi, nthreads);
    return B2R_OK;
}

// ---------------------------------------------------------------- transfers
// A box of global padded coordinates [o, o + n) <-> a dense array (x
// fastest).  Each cell is read from (or, for owned cells, written to) the
// block whose padded region holds it: the owning block for an owned cell, the
// block on the global face for a frame cell.
template <bool TO_FIELD>
__global__ void box_copy_kernel(float* __restrict__ f, float* __restrict__ dense, KGeo g, int ox,
                                int oy, int oz, int nx, int ny, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int c[3] = {ox + (int)(i % nx), oy + (int)((i / nx) % ny), oz + (int)(i / ((long long)nx * ny))};
        int bc[3], lc[3];
        for (int a = 0; a < 3; ++a) {
            int q = c[a] >= 0 ? c[a] / g.B[a] : -1;
            q = q < 0 ? 0 : (q >= g.nb[a] ? g.nb[a] - 1 : q);
            bc[a] = q;
            lc[a] = c[a] - q * g.B[a] + g.h;
        }
        const int b = (bc[2] * g.nb[1] + bc[1]) * g.nb[0] + bc[0];
        const long long e = cell_index(g, b, lc[0], lc[1], lc[2]);
        if (TO_FIELD) f[e] = dense[i];
        else dense[i] = f[e];
    }
}

template <bool TO_FIELD>
static int box_copy(b2r_ctx* ctx, void* f, float* dense, int ox, int oy, int oz, int nx, int ny,
                    int nz)
{
    const long long n = (long long)nx * ny * nz;
    if (n <= 0) return B2R_OK;
    long long nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    box_copy_kernel<TO_FIELD><<<(int)nb, 256, 0, ctx->stream>>>((float*)f, dense, ctx->k, ox, oy,
                                                                 oz, nx, ny, n);
    B2R_LAUNCH_CHECK();
    return B2R_OK;
}

// ---------------------------------------------------------------- hybrid run
extern "C" int b2r_heat_run_hybrid(b2r_ctx* ctx, void* d_a, void* d_b, const b2r_heat_params* p,
                                   int nsteps, int cpu_planes, int nthreads, int* result_in_b,
                                   float* ms)
{
    if (int rc = b2r_heat_check(ctx, p)) return rc;
    if (int rc = b2r_check_field(ctx, d_a)) return rc;
    if (int rc = b2r_check_field(ctx, d_b)) return rc;
    B2R_CHECK(d_a != d_b && nsteps >= 0, B2R_E_ARG, "bad buffers or nsteps");
    KGeo& g = ctx->k;
    B2R_CHECK(!g.periodic && g.fmask == 0x3F && !g.two_d, B2R_E_UNSUPPORTED,
              "hybrid runs need a 3D grid with every global face frozen");
    const int D = p->D, R = ctx->g.R, h = g.h;
    const int X = ctx->g.env_x, Y = ctx->g.env_y, Z = ctx->g.env_z;
    B2R_CHECK(cpu_planes >= h && cpu_planes <= Z - h, B2R_E_ARG,
              "cpu_planes must lie in [R*D, env_z - R*D] = [%d, %d] (got %d)", h, Z - h,
              cpu_planes);
    if (int rc = b2r_activate(ctx)) return rc;
    const int zcut = Z - cpu_planes;  // first plane of the slab
    // host slab: planes zcut-h .. Z+h-1 (bottom joint layer, owned planes,
    // top frame), each (X+2h) x (Y+2h) with its x-y frame
    const int SX = X + 2 * h, SY = Y + 2 * h, SZ = cpu_planes + 2 * h;
    const size_t plane = (size_t)SX * SY, slab = plane * SZ;
    const size_t joint = (size_t)X * Y * h;  // owned cells of h planes
    std::vector<float> sa(slab), sb(slab);
    float* hst = nullptr;   // pinned staging: the whole slab once, then joint layers
    float* dst = nullptr;   // device staging of the same size
    const size_t stage = slab > 2 * joint ? slab : 2 * joint;
    auto release = [&]() {
        g.zclip = 0;
        (void)hipStreamSynchronize(ctx->stream);
        if (hst) (void)hipHostFree(hst);
        if (dst) (void)hipFree(dst);
    };
    if (hipHostMalloc((void**)&hst, stage * 4, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&dst, stage * 4) != hipSuccess) {
        release();
        b2r_set_error("hybrid: staging allocation of %zu MB failed", stage * 4 >> 20);
        return B2R_E_NOMEM;
    }
    int rc = B2R_OK;
    auto hip_ok = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && !rc) {
            b2r_set_error("hybrid: %s: %s", what, hipGetErrorString(e));
            rc = B2R_E_HIP;
        }
    };
    // the slab with its frame, from d_a (both host buffers alike, as d_a and
    // d_b hold the same frame)
    rc = box_copy<false>(ctx, d_a, dst, -h, -h, zcut - h, SX, SY, SZ);
    hip_ok(hipMemcpyAsync(hst, dst, slab * 4, hipMemcpyDeviceToHost, ctx->stream), "slab download");
    hip_ok(hipStreamSynchronize(ctx->stream), "sync");
    if (rc) {
        release();
        return rc;
    }
    memcpy(sa.data(), hst, slab * 4);
    memcpy(sb.data(), hst, slab * 4);

    void* cur = d_a;
    void* nxt = d_b;
    float* scur = sa.data();
    float* snxt = sb.data();
    const size_t row_own = (size_t)X * 4;
    auto t0 = std::chrono::steady_clock::now();
    for (int done = 0; done < nsteps && !rc;) {
        const int n = nsteps - done < R ? nsteps - done : R;
        // joint layers: slab planes zcut .. zcut+h-1 up (host -> device field),
        // the field's planes zcut-h .. zcut-1 down (into the slab)
        float* up = hst;
        float* down = hst + joint;
        for (int k = 0; k < h; ++k)
            for (int y = 0; y < Y; ++y)
                memcpy(up + ((size_t)k * Y + y) * X,
                       scur + (size_t)(h + k) * plane + (size_t)(y + h) * SX + h, row_own);
        hip_ok(hipMemcpyAsync(dst, up, joint * 4, hipMemcpyHostToDevice, ctx->stream), "joint up");
        if (!rc) rc = box_copy<true>(ctx, cur, dst, 0, 0, zcut, X, Y, h);
        if (!rc) rc = box_copy<false>(ctx, cur, dst + joint, 0, 0, zcut - h, X, Y, h);
        hip_ok(hipMemcpyAsync(down, dst + joint, joint * 4, hipMemcpyDeviceToHost, ctx->stream),
               "joint down");
        hip_ok(hipStreamSynchronize(ctx->stream), "sync");
        if (rc) break;
        for (int k = 0; k < h; ++k)
            for (int y = 0; y < Y; ++y)
                memcpy(scur + (size_t)k * plane + (size_t)(y + h) * SX + h,
                       down + ((size_t)k * Y + y) * X, row_own);
        // the GPU part: exchange, then n inner steps with update boxes that
        // end in the joint layer (launched, running while the host computes)
        rc = b2r_direct_exchange(ctx, &cur, 1);
        g.zclip = zcut + h;
        for (int i = 0; i < n && !rc; ++i) {
            rc = b2r_heat_launch(ctx, cur, nxt, p, i);
            void* t = cur;
            cur = nxt;
            nxt = t;
        }
        g.zclip = 0;
        // the slab: n inner steps on the host cores; x and y clipped to the
        // owned cells (frozen faces), z from D(i+1) planes into the joint
        // layer up to the top owned plane
        for (int i = 0; i < n && !rc; ++i) {
            const int lo[3] = {h, h, D * (i + 1)};
            const int hi[3] = {h + X - 1, h + Y - 1, h + cpu_planes - 1};
            cpu_box_any(p, scur, snxt, SX, SY, lo, hi, nthreads);
            float* t = scur;
            scur = snxt;
            snxt = t;
        }
        hip_ok(hipStreamSynchronize(ctx->stream), "sync");
        done += n;
    }
    // the slab's owned planes into the result field
    if (!rc) {
        const size_t own = (size_t)X * Y * cpu_planes;
        float* hs = nullptr;
        hip_ok(hipHostMalloc((void**)&hs, own * 4, hipHostMallocDefault), "host staging");
        float* ds = nullptr;
        hip_ok(hipMalloc((void**)&ds, own * 4), "device staging");
        if (!rc) {
            for (int k = 0; k < cpu_planes; ++k)
                for (int y = 0; y < Y; ++y)
                    memcpy(hs + ((size_t)k * Y + y) * X,
                           scur + (size_t)(h + k) * plane + (size_t)(y + h) * SX + h, row_own);
            hip_ok(hipMemcpyAsync(ds, hs, own * 4, hipMemcpyHostToDevice, ctx->stream), "slab up");
            if (!rc) rc = box_copy<true>(ctx, cur, ds, 0, 0, zcut, X, Y, cpu_planes);
            hip_ok(hipStreamSynchronize(ctx->stream), "sync");
        }
        if (hs) (void)hipHostFree(hs);
        if (ds) (void)hipFree(ds);
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    release();
    if (rc) return rc;
    if (result_in_b) *result_in_b = cur == d_b ? 1 : 0;
    if (ms) *ms = (float)(secs * 1e3);
    return B2R_OK;
}
