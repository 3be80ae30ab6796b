// b2r_tune.hip -- autotuner for launch shapes.
//
// No counterpart in the reference (its CUDA_BLOCK_SIZE is a compile-time 32 or
// 16, heat-3D.cu:8, eqwp.cu:25).  For one kernel kind on the context's actual
// geometry it times every candidate shape with back-to-back launches between
// HIP events (b2r_time_steps), keeps the fastest in the context and appends one
// line per search to a plain-text record.  The fused Game of Life candidates
// are timed with the exchange that precedes each launch in b2r_gol_run, since
// the bit-sliced kernel reads its halo from the neighbouring blocks and needs
// none while the byte-lane kernel does (kernel time alone would not show it):
//   kind=<k> S=<Sx>x<Sy>x<Sz> blocks=<n> R=<R> radius=<r> best=<field>:<v>,... ms=<t> tried=<n>
#include "b2r_common.h"

#include <cstdio>
#include <cstring>
#include <utility>
#include <vector>

extern "C" int b2r_time_steps(b2r_ctx* ctx, int kind, void* d_a, void* d_b,
                              const b2r_heat_params* p, int iters, float* ms);

// iters windows of exchange + fused Game of Life launch, as b2r_gol_run runs
// them (on the context's packed pair, packed from d_a first, where it keeps
// the fields packed)
static int gol_windows(b2r_ctx* ctx, void* d_a, void* d_b, int iters, float* ms)
{
    if (int rc = b2r_activate(ctx)) return rc;
    const int packed = b2r_gol_fused_packed(ctx);
    if (packed)
        if (int rc = b2r_gol_pack_pair(ctx, d_a, &d_a, &d_b)) return rc;
    if (int rc = b2r_time_begin(ctx, ms)) return rc;
    const bool pulls = b2r_gol_fused_reads_halo(ctx);
    for (int it = 0; it < iters; ++it) {
        void* in = (it & 1) ? d_b : d_a;
        void* out = (it & 1) ? d_a : d_b;
        if (ctx->tune.gol_exchange) {
            if (int rc = b2r_boundary_updating(ctx, it, in)) return rc;
        } else if (!pulls) {
            if (int rc = b2r_direct_exchange(ctx, &in, 1)) return rc;
        }
        if (int rc = b2r_gol_fused_launch(ctx, in, out, packed)) return rc;
    }
    return b2r_time_end(ctx, ms);
}

extern "C" int b2r_autotune(b2r_ctx* ctx, int kind, void* d_a, void* d_b,
                            const b2r_heat_params* p, const char* record_path, b2r_tuning* out)
{
    B2R_CHECK(ctx && d_a && d_b && kind >= 0 && kind <= 3, B2R_E_ARG, "bad argument");
    const b2r_tuning base = ctx->tune;
    std::vector<b2r_tuning> cand;
    const int zcs[] = {0, 8, 16, 32, 64, 128, 256};
    // heat also tries chunk lengths between the powers of two, whose workgroup
    // counts can fill the CUs' slots in whole rounds where no power of two does
    const int zcs_heat[] = {0, 8, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 768};
    if (kind == 0 || kind == 3) {
        const int rys[] = {1, 2, 4, 8};
        const int wys[] = {4, 8, 16};
        if (kind == 3) {  // the fused kernel's tile is fixed; only its z chunk varies
            for (int zc : zcs_heat) {
                b2r_tuning t = base;
                t.heat_zc = zc;
                cand.push_back(t);
            }
        } else {
            std::vector<std::pair<int, int>> shapes;  // (heat_wy, heat_wx)
            for (int wy : wys)
                for (int wx = 1; wx * wy <= 16; wx *= 2) shapes.push_back({wy, wx});
            shapes.push_back({4, 3});
            shapes.push_back({2, 6});
            shapes.push_back({3, 3});
            for (const auto& sh : shapes)
                for (int ry : rys)
                    for (int zc : zcs_heat) {
                        b2r_tuning t = base;
                        t.heat_ry = ry;
                        t.heat_wy = sh.first;
                        t.heat_wx = sh.second;
                        t.heat_zc = zc;
                        cand.push_back(t);
                    }
        }
    } else {
        for (int tile = 1; tile <= (kind == 2 ? 15 : 1); ++tile)
            for (int zc : zcs) {
                if (kind == 2 && !b2r_gol_tile_ok(ctx, tile)) continue;
                b2r_tuning t = base;
                if (kind == 1) {
                    t.gol_step_zc = zc;
                } else {
                    t.gol_tile = tile;
                    t.gol_zc = zc;
                }
                cand.push_back(t);
            }
    }
    // fused Game of Life windows last about 0.1 ms: more of them keep the
    // choice stable from run to run
    const int warm = 2, iters = kind == 2 ? 24 : 6;
    float best_ms = 1e30f;
    b2r_tuning best = base;
    int tried = 0;
    for (const b2r_tuning& t : cand) {
        ctx->tune = t;
        float ms = 0;
        int rc;
        if (kind == 2) {
            rc = gol_windows(ctx, d_a, d_b, warm, &ms);
            if (!rc) rc = gol_windows(ctx, d_a, d_b, iters, &ms);
        } else {
            rc = b2r_time_steps(ctx, kind, d_a, d_b, p, warm, &ms);
            if (!rc) rc = b2r_time_steps(ctx, kind, d_a, d_b, p, iters, &ms);
        }
        if (rc) {
            ctx->tune = base;
            return rc;
        }
        ++tried;
        if (ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    }
    ctx->tune = best;
    // a fused Game of Life choice keeps its z chunk explicit (the automatic
    // chunk it stood for on this geometry)
    if (kind == 2 && best.gol_zc == 0) ctx->tune.gol_zc = best.gol_zc = b2r_gol_fused_chunk(ctx);
    if (out) *out = best;
    if (record_path && *record_path) {
        FILE* f = fopen(record_path, "a");
        B2R_CHECK(f, B2R_E_ARG, "cannot open %s", record_path);
        static const char* names[] = {"heat_step", "gol_step", "gol_fused", "heat_fused"};
        fprintf(f, "kind=%s S=%dx%dx%d blocks=%d R=%d radius=%d", names[kind], ctx->L.S[0],
                ctx->L.S[1], ctx->L.S[2], ctx->L.nblocks, ctx->g.R, ctx->g.radius);
        if (kind == 0) {
            fprintf(f, " heat_D=%d best=heat_wy:%d,heat_wx:%d,heat_ry:%d,heat_zc:%d", p ? p->D : 0,
                    best.heat_wy, best.heat_wx, best.heat_ry, best.heat_zc);
        } else if (kind == 3) {
            fprintf(f, " heat_D=%d best=heat_zc:%d", p ? p->D : 0, best.heat_zc);
        } else if (kind == 1) {
            fprintf(f, " best=gol_step_zc:%d", best.gol_step_zc);
        } else {
            fprintf(f, " best=gol_tile:%d,gol_zc:%d", best.gol_tile, best.gol_zc);
        }
        fprintf(f, " %s=%.5f tried=%d\n", kind == 2 ? "ms_per_window" : "ms_per_launch",
                best_ms / iters, tried);
        fclose(f);
    }
    return B2R_OK;
}
