// A fixed batch of independent subdomain solves, prepared once and replayed:
// the batched workload of many PoissonSolver2D objects of uneven sizes
// (the reference solves them one after another,
// pe/poisson/poisson_solver2d.cpp:62-90).
//
// fp_sub_solve_batched builds its launch descriptors on every call and, for
// more than 8 problems, copies them to the device through pinned staging:
// host work per call, and nothing a HIP graph can record.  Here the
// descriptors of every launch (forward transforms, x sweeps, inverse
// transforms) are built once, for fields bound at creation, and copied to
// one device buffer that the batch owns; a solve is then a fixed sequence of
// kernel launches that read those tables, so it can be captured and
// replayed.  Two choices differ from the per-call path:
//   * one transform shape for every problem where it fits (128 threads per
//     workgroup, three pairs per odd-prime work item), so that problems of
//     different lengths and different tuned shapes share launches: one launch
//     per phase instead of one per tuned shape;
//   * within each launch the problems are ordered by decreasing work per
//     workgroup (FFT length, sweep length): the dispatcher issues workgroups
//     in order, so the longest run first and the short ones fill the tail
//     (longest-processing-time order).
#include "fp_internal.h"

#include <cstdio>
#include <cstdlib>
#include <memory>

namespace fp {

struct BatchBase {
    virtual ~BatchBase() = default;
    virtual void solve(hipStream_t s) = 0;
};

template <typename T>
static const TransformPlan<T>* plan_of(const SubPlan& sp)
{
    if constexpr (sizeof(T) == 8) return &sp.t64;
    else return &sp.t32;
}

template <typename T>
struct Batch : BatchBase {
    std::vector<GroupPlan> fw, tri, iv;
    std::vector<size_t>    off_fw, off_tri, off_iv;
    DevBuf<unsigned char>  tables;

    void build(int nprob, fp_sub_s* const* subs, void* const* d_f, const int* nrhs)
    {
        std::vector<TransformJob<T>> tj;
        std::vector<TridiagJob<T>>   rj;
        // the common transform shape; FP_BATCH_SHAPE=<threads>,<form>
        // (measurements) overrides it, {0, -1} keeping every plan's own
        std::array<int, 2> shape = {128, 3};
        if (const char* e = std::getenv("FP_BATCH_SHAPE")) {
            int a = 0, b = -1;
            if (std::sscanf(e, "%d,%d", &a, &b) == 2) shape = {a, b};
        }
        for (int k = 0; k < nprob; k++) {
            if (nrhs[k] <= 0) continue;
            SubPlan& sp = subs[k]->plan;
            if constexpr (sizeof(T) == 4) sp.ensure_f32();
            TransformJob<T> j{};
            j.plan  = plan_of<T>(sp);
            j.data  = d_f[k];
            j.ld    = (long long)sp.nx * sp.ny;
            j.nx    = sp.nx;
            j.nrhs  = nrhs[k];
            j.dst64 = sizeof(T) == 8;
            apply_batch_shape(j, shape);
            tj.push_back(j);
            rj.push_back({&sp, static_cast<T*>(d_f[k]), (long long)sp.nx * sp.ny, nrhs[k], 1});
        }
        fw  = plan_ytransform_groups<T>(tj.data(), (int)tj.size(), 0);
        tri = plan_tridiag_groups<T>(rj.data(), (int)rj.size());
        iv  = plan_ytransform_groups<T>(tj.data(), (int)tj.size(), 1);
        // one device buffer for every table, each at a 256-byte boundary
        size_t total = 0;
        auto   place = [&](const std::vector<GroupPlan>& gs, std::vector<size_t>& off) {
            for (const GroupPlan& g : gs) {
                off.push_back(total);
                total = (total + g.table.size() + 255) & ~(size_t)255;
            }
        };
        place(fw, off_fw);
        place(tri, off_tri);
        place(iv, off_iv);
        if (total == 0) return;
        std::vector<unsigned char> h(total, 0);
        auto copy = [&](const std::vector<GroupPlan>& gs, const std::vector<size_t>& off) {
            for (size_t g = 0; g < gs.size(); g++) std::memcpy(h.data() + off[g], gs[g].table.data(), gs[g].table.size());
        };
        copy(fw, off_fw);
        copy(tri, off_tri);
        copy(iv, off_iv);
        tables.alloc(total);
        tables.upload(h.data(), total);
    }

    void solve(hipStream_t s) override
    {
        for (size_t g = 0; g < fw.size(); g++) launch_ytransform_group<T>(fw[g], tables.p + off_fw[g], s);
        for (size_t g = 0; g < tri.size(); g++) launch_tridiag_group<T>(tri[g], tables.p + off_tri[g], s);
        for (size_t g = 0; g < iv.size(); g++) launch_ytransform_group<T>(iv[g], tables.p + off_iv[g], s);
    }
};

}  // namespace fp

struct fp_batch_s {
    fp_ctx_s*                      ctx = nullptr;
    std::unique_ptr<fp::BatchBase> impl;
};

extern "C" {

int fp_batch_create(fp_ctx ctx, int nprob, const fp_sub* subs, void* const* d_f, const int* nrhs, int fp32,
                    fp_batch* out)
{
    return fp::guarded([&] {
        FP_REQUIRE(ctx && out && nprob >= 0 && (nprob == 0 || (subs && d_f && nrhs)), "null argument");
        FP_HIP(hipSetDevice(ctx->device));
        for (int k = 0; k < nprob; k++) {
            FP_REQUIRE(subs[k] && nrhs[k] >= 0 && (nrhs[k] == 0 || d_f[k]), "null problem");
            FP_REQUIRE(subs[k]->ctx->device == ctx->device, "a subdomain of another device");
        }
        auto b = std::make_unique<fp_batch_s>();
        b->ctx = ctx;
        if (fp32) {
            auto impl = std::make_unique<fp::Batch<float>>();
            impl->build(nprob, subs, d_f, nrhs);
            b->impl = std::move(impl);
        } else {
            auto impl = std::make_unique<fp::Batch<double>>();
            impl->build(nprob, subs, d_f, nrhs);
            b->impl = std::move(impl);
        }
        *out = b.release();
    });
}

int fp_batch_solve(fp_batch b, void* stream)
{
    return fp::guarded([&] {
        FP_REQUIRE(b && b->impl, "null batch");
        FP_HIP(hipSetDevice(b->ctx->device));
        b->impl->solve((hipStream_t)stream);
    });
}

void fp_batch_destroy(fp_batch b) { delete b; }

}  // extern "C"
