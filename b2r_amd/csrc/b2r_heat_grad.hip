// b2r_heat_grad.hip -- hand-written discrete adjoint of the heat time loop.
//
// No reference counterpart: the reference only runs the heat model forward
// (heat_3D_gpu_main heat-3D.cu:583-642).  Misfit and parameters:
//   J = 1/2 sum_owned (u_N - target)^2,   u_{n+1} = A u_n,
//   A = I + sum_a k_a L_a     (F32 mode; REF mode has k_a = adt / h2_a),
// with L_a the second difference of fdm_heat_diffuse_deltaD along axis a.
// With frozen (zero-Dirichlet for the adjoint) or periodic global faces every
// L_a restricted to the owned cells is symmetric, so A^T = A and the adjoint
// recursion is the forward step itself applied to a zero-framed field:
//   lam_N = u_N - target,  lam_n = A lam_{n+1},
//   dJ/du_0 = lam_0,       dJ/dk_a = sum_n <lam_{n+1}, L_a u_n>.
// The forward states come from checkpoints every ckpt steps; each segment is
// recomputed (its states kept on the device) before its reverse sweep.  The
// dot products and the misfit are reduced in FP64 (wave64 shuffles, then a
// fixed-order sum, so the result is deterministic).
#include "b2r_common.h"
#include "b2r_heat_ops.h"

#include <vector>

constexpr int HG_TX = 64, HG_TY = 4;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// owned-cell tile of the workgroup: 64 x 4 columns of one block, z marched
struct OwnedTile {
    int b, x, y;
    bool in;
};

__device__ __forceinline__ OwnedTile owned_tile(const KGeo& g, int ntx, int nty)
{
    int t = blockIdx.x;
    const int txi = t % ntx;
    t /= ntx;
    const int tyi = t % nty;
    OwnedTile o;
    o.b = t / nty;
    const int lx = txi * HG_TX + threadIdx.x, ly = tyi * HG_TY + threadIdx.y;
    o.in = lx < g.B[0] && ly < g.B[1];
    o.x = g.h + lx;
    o.y = g.h + ly;
    return o;
}

// per-workgroup sums of NC values -> part[blockIdx.x * NC + a]
template <int NC>
__device__ __forceinline__ void block_sums(double (&v)[NC], double* __restrict__ part)
{
    __shared__ double ws[NC][HG_TY];
    const int w = threadIdx.y;  // one wave per row of the 64 x 4 workgroup
#pragma unroll
    for (int a = 0; a < NC; ++a) {
        const double s = wave_sum(v[a]);
        if (threadIdx.x == 0) ws[a][w] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0 && w == 0) {
#pragma unroll
        for (int a = 0; a < NC; ++a)
            part[(long long)blockIdx.x * NC + a] = (ws[a][0] + ws[a][1]) + (ws[a][2] + ws[a][3]);
    }
}

// lam = u - target on owned cells; partial sums of (u - target)^2
__global__ __launch_bounds__(256) void heat_resid_kernel(const float* __restrict__ u,
                                                         const float* __restrict__ tgt,
                                                         float* __restrict__ lam, KGeo g, int ntx,
                                                         int nty, double* __restrict__ part)
{
    const OwnedTile o = owned_tile(g, ntx, nty);
    double acc[1] = {0.0};
    if (o.in) {
        for (int z = g.h; z < g.h + g.B[2]; ++z) {
            const long long c = cell_index(g, o.b, o.x, o.y, z);
            const float r = u[c] - tgt[c];
            lam[c] = r;
            acc[0] += (double)r * (double)r;
        }
    }
    block_sums<1>(acc, part);
}

// partial sums of lam[c] * L_a u[c], a = x, y, z, over owned cells
template <int D, int MODE>
__global__ __launch_bounds__(256) void heat_kdot_kernel(const float* __restrict__ lam,
                                                        const float* __restrict__ u, KGeo g,
                                                        int ntx, int nty, double* __restrict__ part)
{
    const OwnedTile o = owned_tile(g, ntx, nty);
    double acc[3] = {0.0, 0.0, 0.0};
    if (o.in) {
        const long long sy = g.pitch, sz = g.slice;
        for (int z = g.h; z < g.h + g.B[2]; ++z) {
            const long long c = cell_index(g, o.b, o.x, o.y, z);
            const float uc = u[c];
            float mx[D + 1], px[D + 1], my[D + 1], py[D + 1], mz[D + 1], pz[D + 1];
#pragma unroll
            for (int j = 1; j <= D; ++j) {
                mx[j] = u[c - j];
                px[j] = u[c + j];
                my[j] = u[c - j * sy];
                py[j] = u[c + j * sy];
                mz[j] = u[c - j * sz];
                pz[j] = u[c + j * sz];
            }
            const double l = (double)lam[c];
            acc[0] += l * (double)lap<D, MODE>(mx, px, uc);
            acc[1] += l * (double)lap<D, MODE>(my, py, uc);
            acc[2] += l * (double)lap<D, MODE>(mz, pz, uc);
        }
    }
    block_sums<3>(acc, part);
}

// acc[a] += sum_i part[i * nc + a] in a fixed order (one workgroup)
__global__ __launch_bounds__(256) void heat_psum_kernel(const double* __restrict__ part, int n,
                                                        int nc, double* __restrict__ acc)
{
    __shared__ double ws[4];
    const int tid = threadIdx.x;
    for (int a = 0; a < nc; ++a) {
        double s = 0.0;
        for (int i = tid; i < n; i += 256) s += part[(long long)i * nc + a];
        s = wave_sum(s);
        if ((tid & 63) == 0) ws[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) acc[a] += (ws[0] + ws[1]) + (ws[2] + ws[3]);
        __syncthreads();
    }
}

template <int D, int MODE>
static void launch_kdot(b2r_ctx* ctx, const float* lam, const float* u, int ntx, int nty,
                        double* part)
{
    const int nwg = ntx * nty * ctx->k.nblk;
    heat_kdot_kernel<D, MODE><<<nwg, dim3(HG_TX, HG_TY), 0, ctx->stream>>>(lam, u, ctx->k, ntx,
                                                                           nty, part);
}

static void kdot(b2r_ctx* ctx, const b2r_heat_params* p, const float* lam, const float* u,
                 int ntx, int nty, double* part)
{
#define B2R_KDOT_CASE(DD)                                                             \
    case DD:                                                                          \
        if (p->mode == B2R_HEAT_F32) launch_kdot<DD, B2R_HEAT_F32>(ctx, lam, u, ntx, nty, part); \
        else launch_kdot<DD, B2R_HEAT_REF>(ctx, lam, u, ntx, nty, part);             \
        break;
    switch (p->D) {
        B2R_KDOT_CASE(1)
        B2R_KDOT_CASE(2)
        B2R_KDOT_CASE(3)
        B2R_KDOT_CASE(4)
    }
#undef B2R_KDOT_CASE
}

extern "C" int b2r_heat_gradient(b2r_ctx* ctx, const void* d_u0, const void* d_target,
                                 const b2r_heat_params* p, int nsteps, int ckpt, double* misfit,
                                 void* d_grad_u0, double* grad_k)
{
    if (int rc = b2r_heat_check(ctx, p)) return rc;
    B2R_CHECK(misfit && grad_k && nsteps >= 1 && ckpt >= 1, B2R_E_ARG,
              "bad argument (nsteps=%d, ckpt=%d)", nsteps, ckpt);
    for (const void* f : {d_u0, d_target, (const void*)d_grad_u0})
        if (int rc = b2r_check_field(ctx, f)) return rc;
    B2R_CHECK(closed_faces(ctx->k), B2R_E_UNSUPPORTED,
              "the heat adjoint needs periodic or frozen global faces");
    if (int rc = b2r_activate(ctx)) return rc;
    const KGeo& g = ctx->k;
    const int ntx = ceil_div(g.B[0], HG_TX), nty = ceil_div(g.B[1], HG_TY);
    const int nwg = ntx * nty * g.nblk;
    const int ncp = (nsteps + ckpt - 1) / ckpt;

    std::vector<void*> fields;
    double* dbuf = nullptr;  // [acc 4][part 3 * nwg]
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(ctx->stream);
        for (void* f : fields) b2r_field_free(ctx, f);
        if (dbuf) (void)hipFree(dbuf);
    };
    auto new_field = [&](void** f) {
        const int rc = b2r_field_alloc(ctx, f);
        if (rc == B2R_OK) fields.push_back(*f);
        return rc;
    };
    // checkpoints, segment states (each carries u0's frozen frame), adjoint pair
    std::vector<void*> cp(ncp), seg(std::max(ckpt, 2));
    void* lam[2] = {nullptr, nullptr};
    int rc = B2R_OK;
    for (auto& f : cp)
        if (!rc) rc = new_field(&f);
    for (auto& f : seg)
        if (!rc) rc = new_field(&f);
    for (auto& f : lam)
        if (!rc) rc = new_field(&f);
    if (!rc && hipMalloc(&dbuf, sizeof(double) * (4 + 3 * (size_t)nwg)) != hipSuccess) {
        dbuf = nullptr;
        b2r_set_error("heat gradient: workspace allocation failed");
        rc = B2R_E_NOMEM;
    }
    if (rc) {
        cleanup();
        return rc;
    }
    double* acc = dbuf;
    double* part = dbuf + 4;
    const float zero = 0.0f;
    if (hipMemsetAsync(acc, 0, 4 * sizeof(double), ctx->stream) != hipSuccess) {
        b2r_set_error("heat gradient: workspace clear failed");
        rc = B2R_E_HIP;
    }
    for (void* f : seg)
        if (!rc) rc = b2r_field_copy(ctx, f, d_u0);
    for (void* f : lam)
        if (!rc) rc = b2r_field_fill(ctx, f, &zero);
    if (!rc) rc = b2r_field_copy(ctx, cp[0], d_u0);

    // forward, checkpointing the state at the start of every segment
    void* cur = seg[0];
    void* nxt = seg[1];
    for (int j = 0; j < ncp && !rc; ++j) {
        if (j > 0) rc = b2r_field_copy(ctx, cp[j], cur);
        const int n = std::min(ckpt, nsteps - j * ckpt);
        int in_b = 0;
        if (!rc) rc = b2r_heat_run(ctx, cur, nxt, p, n, 1, &in_b, nullptr);
        if (!rc && in_b) std::swap(cur, nxt);
    }
    const dim3 blk(HG_TX, HG_TY);
    if (!rc) {
        heat_resid_kernel<<<nwg, blk, 0, ctx->stream>>>((const float*)cur, (const float*)d_target,
                                                        (float*)lam[0], g, ntx, nty, part);
        heat_psum_kernel<<<1, 256, 0, ctx->stream>>>(part, nwg, 1, acc + 3);
        rc = hipGetLastError() == hipSuccess ? B2R_OK : B2R_E_HIP;
        if (rc) b2r_set_error("heat gradient: residual launch failed");
    }

    // reverse sweep, segment by segment
    void* la = lam[0];
    void* lb = lam[1];
    for (int j = ncp - 1; j >= 0 && !rc; --j) {
        const int t0 = j * ckpt, n = std::min(ckpt, nsteps - t0);
        rc = b2r_field_copy(ctx, seg[0], cp[j]);
        // seg[k] = u_{t0+k}, halos refreshed (the stencil of the dot reads them)
        for (int k = 0; k < n && !rc; ++k) {
            rc = b2r_direct_exchange(ctx, &seg[k], 1);
            if (!rc && k + 1 < n) rc = b2r_heat_launch(ctx, seg[k], seg[k + 1], p, 0);
        }
        for (int k = n - 1; k >= 0 && !rc; --k) {
            kdot(ctx, p, (const float*)la, (const float*)seg[k], ntx, nty, part);
            heat_psum_kernel<<<1, 256, 0, ctx->stream>>>(part, nwg, 3, acc);
            if (hipGetLastError() != hipSuccess) {
                b2r_set_error("heat gradient: dot launch failed");
                rc = B2R_E_HIP;
            }
            if (!rc) rc = b2r_direct_exchange(ctx, &la, 1);
            if (!rc) rc = b2r_heat_launch(ctx, la, lb, p, 0);
            std::swap(la, lb);
        }
    }
    double h_acc[4] = {0, 0, 0, 0};
    if (!rc) rc = b2r_field_copy(ctx, d_grad_u0, la);
    if (!rc && hipMemcpyAsync(h_acc, acc, sizeof(h_acc), hipMemcpyDeviceToHost, ctx->stream) !=
                   hipSuccess) {
        b2r_set_error("heat gradient: result copy failed");
        rc = B2R_E_HIP;
    }
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) {
        b2r_set_error("heat gradient: %s", hipGetErrorString(hipGetLastError()));
        rc = B2R_E_HIP;
    }
    if (!rc) {
        *misfit = 0.5 * h_acc[3];
        for (int a = 0; a < 3; ++a) grad_k[a] = h_acc[a];
    }
    cleanup();
    return rc;
}
