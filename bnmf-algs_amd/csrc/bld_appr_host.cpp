// bld_appr_host.cpp — host side of the approximate BLD solver bld::bld_appr
// (src/bld/bld_appr.hpp:287-372): planning, HBM layout and the captured
// per-iteration launch sequence of the kernels in bld_appr.hip.
//
// The reference materialises S, grad_plus and four x*y*z loops per
// half-step on the host. Here the factors nu (x*z) and mu (z*y) stay on the
// device, S is formed per fiber inside each pass from X and the factors, and
// an iteration is, per half-step: sums pass, prep + psi(A) (shared with the
// multiplicative solver), step pass, factor update — captured once into a
// hipGraph and replayed.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "bnmf_internal.hpp"

namespace bnmf {

namespace {
constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Rows per tile: enough tiles to give an instance's launch about one block
// per 2048 fibers, between 16 and 512 blocks, with tiles of 4 to 64 rows.
// The plan depends on the shape only, so an instance gives the same bits
// alone and batched. Measured on MI355X: 1024x1024x32 alone, 512 blocks
// (32-row tiles) 314 us per iteration against 518 us at 256 or fewer, and
// 2048 or 8192 blocks slower again (more tiles add column partials that the
// factor and prep kernels must reduce); the 256-instance batch (x, y up to
// 512), 3524 us per iteration at 512 blocks per instance, 2706 us at 128,
// 2566 us at 16 (the factor kernel's partial sums over row tiles dominate
// once every instance has many short tiles). BNMF_APPR_BLOCKS sets a fixed
// target for every instance instead, for experiments.
uint32_t appr_ti_rows(uint32_t x, uint32_t y, uint32_t n_tj) {
    static const uint32_t fixed_target = [] {
        const char* e = std::getenv("BNMF_APPR_BLOCKS");
        const long v = e ? std::strtol(e, nullptr, 10) : 0L;
        return v > 0 ? (uint32_t)std::min<long>(v, 1 << 20) : 0u;
    }();
    const uint64_t by_size = ((uint64_t)x * y) / 2048u;
    const uint32_t target =
        fixed_target ? fixed_target : (uint32_t)std::min<uint64_t>(std::max<uint64_t>(by_size, 16u), 512u);
    const uint32_t want = std::max<uint32_t>(1, target / std::max<uint32_t>(n_tj, 1));
    uint32_t ti = ceil_div(x, want);
    ti = (ti + 3) / 4 * 4;
    return std::min<uint32_t>(std::max<uint32_t>(ti, 4), 64);
}
}  // namespace

ApprBatch::ApprBatch(int dtype, double eps, const std::vector<ApprSpec>& specs)
    : dtype_(dtype), es_(dtype == BNMF_F32 ? 4 : 8) {
    ensure_runtime();
    const int psi = dtype == BNMF_F32 ? PSIM_F32 : PSIM_F64_APPR;
    psi_ = psi;
    // layouts
    size_t total = 0;
    for (const ApprSpec& s : specs) {
        BNMF_REQUIRE(s.z >= 1, "bld_appr needs z >= 1");
        L_ l{};
        l.x = s.x;
        l.y = s.y;
        l.z = s.z;
        // fibers deeper than 64 take the wavefront-per-fiber pass on
        // kWideRows x kWideCols tiles
        const bool deep = s.z > kFusedZ;
        l.n_tj = ceil_div(s.y, deep ? kWideCols : kTileCols);
        l.ti_rows = deep ? kWideRows : appr_ti_rows(s.x, s.y, l.n_tj);
        l.n_ti = ceil_div(s.x, l.ti_rows);
        l.nA = ceil_div(s.x, kPrepRows);
        l.n_cj = ceil_div(s.y, kPrepCols);
        size_t off = total;
        auto take = [&](size_t n) {
            const size_t o = off;
            off = align_up(off + std::max<size_t>(n, 1));
            return o;
        };
        const size_t x = s.x, y = s.y, z = s.z;
        l.off_X = take(x * y * es_);
        l.off_nu = take(x * z * es_);
        l.off_mu = take(z * y * es_);
        l.off_nu0 = take(x * z * es_);
        l.off_mu0 = take(z * y * es_);
        l.off_alpha = take(x * 8);
        l.off_beta = take(z * 8);
        l.off_psib = take(y * z * es_);
        l.off_psia = take(x * z * es_);
        l.off_partA = take((size_t)l.nA * z * es_);
        l.off_psiA = take(z * es_);
        l.off_pjk = take((size_t)l.n_ti * y * z * es_);
        l.off_pjk2 = take((size_t)l.n_ti * y * z * es_);
        l.off_ipk = take((size_t)l.n_tj * x * z * es_);
        l.off_ipk2 = take((size_t)l.n_tj * x * z * es_);
        l.off_S = take(x * y * z * es_);
        total = off;
        lay_.push_back(l);
    }
    arena_ = DeviceBuffer(std::max<size_t>(total, kAlign));

    // uploads; nu0 / mu0 drawn when not given: rows of nu and columns of mu
    // are Dirichlet(1_z) (bld_appr.hpp:302-318), from the init_S stream of
    // `seed` over x + y unit fibers (rows of nu first, then columns of mu)
    std::vector<ApprInst> ai;
    std::vector<BldInst> bi;
    std::vector<char> tmp;
    for (size_t n = 0; n < specs.size(); ++n) {
        const ApprSpec& s = specs[n];
        const L_& l = lay_[n];
        char* b = arena_.as<char>();
        const size_t x = s.x, y = s.y, z = s.z;
        if (x * y) BNMF_HIP(hipMemcpy(b + l.off_X, s.X, x * y * es_, hipMemcpyHostToDevice));
        if (x) BNMF_HIP(hipMemcpy(b + l.off_alpha, s.alpha, x * 8, hipMemcpyHostToDevice));
        BNMF_HIP(hipMemcpy(b + l.off_beta, s.beta, z * 8, hipMemcpyHostToDevice));
        const void* nu0 = s.nu0;
        const void* mu0 = s.mu0;
        std::vector<char> dn, dm;
        if ((!nu0 || !mu0) && x + y > 0) {
            tmp.resize((x + y) * z * es_);
            if (dtype == BNMF_F32) {
                std::vector<float> ones(x + y, 1.0f);
                init_S<float>(s.seed, ones.data(), x + y, 1, z, reinterpret_cast<float*>(tmp.data()));
            } else {
                std::vector<double> ones(x + y, 1.0);
                init_S<double>(s.seed, ones.data(), x + y, 1, z, reinterpret_cast<double*>(tmp.data()));
            }
            if (!nu0) {
                dn.assign(tmp.data(), tmp.data() + x * z * es_);
                nu0 = dn.data();
            }
            if (!mu0) {
                // columns of mu: transpose the y rows drawn after nu's
                dm.resize(z * y * es_);
                for (size_t j = 0; j < y; ++j)
                    for (size_t k = 0; k < z; ++k)
                        std::memcpy(dm.data() + (k * y + j) * es_, tmp.data() + ((x + j) * z + k) * es_, es_);
                mu0 = dm.data();
            }
        }
        if (x * z) BNMF_HIP(hipMemcpy(b + l.off_nu0, nu0, x * z * es_, hipMemcpyHostToDevice));
        if (z * y) BNMF_HIP(hipMemcpy(b + l.off_mu0, mu0, z * y * es_, hipMemcpyHostToDevice));

        ApprInst a{};
        a.X = b + l.off_X;
        a.nu = b + l.off_nu;
        a.mu = b + l.off_mu;
        a.psib = b + l.off_psib;
        a.psia = b + l.off_psia;
        a.psiA = b + l.off_psiA;
        a.part_pjk = b + l.off_pjk;
        a.part_ipk = b + l.off_ipk;
        a.part_pjk2 = b + l.off_pjk2;
        a.part_ipk2 = b + l.off_ipk2;
        a.S = b + l.off_S;
        a.eps = eps;
        a.x = s.x;
        a.y = s.y;
        a.z = s.z;
        a.ti_rows = l.ti_rows;
        a.n_ti = l.n_ti;
        a.n_tj = l.n_tj;
        uint32_t zq = 1, lg = 0;
        while (zq < s.z) {
            zq <<= 1;
            ++lg;
        }
        a.zq = zq;
        a.zq_log2 = lg;
        ai.push_back(a);

        // the prep view of the same buffers
        BldInst d{};
        d.alpha = reinterpret_cast<const double*>(b + l.off_alpha);
        d.beta = reinterpret_cast<const double*>(b + l.off_beta);
        d.psib = b + l.off_psib;
        d.psia = b + l.off_psia;
        d.partA = b + l.off_partA;
        d.psiA = b + l.off_psiA;
        d.part_pjk = b + l.off_pjk;
        d.part_ipk = b + l.off_ipk;
        d.eps = eps;
        d.inv_z = 1.0f / (float)s.z;
        d.x = s.x;
        d.y = s.y;
        d.z = s.z;
        d.n_ti = l.n_ti;
        d.n_tj = l.n_tj;
        d.nA = l.nA;
        d.n_cj = l.n_cj;
        d.zq = zq;
        d.zq_log2 = lg;
        bi.push_back(d);
    }

    // work lists: pass items by z bucket, prep items, factor items
    std::map<uint32_t, std::vector<PassItem>> by_zb;
    std::map<uint32_t, uint32_t> zmax;
    std::vector<PrepItem> prep;
    std::vector<PassItem> fnu, fmu;
    for (size_t n = 0; n < specs.size(); ++n) {
        const L_& l = lay_[n];
        if ((size_t)l.x * l.y == 0) continue;
        const uint32_t zb = z_bucket(l.z);
        auto& v = by_zb[zb];
        for (uint32_t ti = 0; ti < l.n_ti; ++ti)
            for (uint32_t tj = 0; tj < l.n_tj; ++tj) v.push_back(PassItem{(uint32_t)n, ti, tj, 0});
        zmax[zb] = std::max(zmax[zb], l.z);
        for (uint32_t c = 0; c < l.n_cj; ++c) prep.push_back(PrepItem{(uint32_t)n, 0, c, 0});
        for (uint32_t c = 0; c < l.nA; ++c) prep.push_back(PrepItem{(uint32_t)n, 1, c, 0});
        // factor updates: four rows (columns) per item, one per wavefront
        for (uint32_t c = 0; c < ceil_div(l.x, 4); ++c) fnu.push_back(PassItem{(uint32_t)n, c, 0, 0});
        for (uint32_t c = 0; c < ceil_div(l.y, 4); ++c) fmu.push_back(PassItem{(uint32_t)n, c, 0, 0});
    }
    size_t moff = 0;
    auto mtake = [&](size_t nbytes) {
        const size_t o = moff;
        moff = align_up(moff + std::max<size_t>(nbytes, 4));
        return o;
    };
    off_ai_ = mtake(ai.size() * sizeof(ApprInst));
    off_bi_ = mtake(bi.size() * sizeof(BldInst));
    for (auto& kv : by_zb) {
        Group g;
        g.zb = kv.first;
        g.zmax = zmax[kv.first];
        g.n_items = (uint32_t)kv.second.size();
        g.off_items = mtake(kv.second.size() * sizeof(PassItem));
        g.off_count = mtake(4);
        groups_.push_back(g);
    }
    n_prep_ = (uint32_t)prep.size();
    off_prep_ = mtake(prep.size() * sizeof(PrepItem));
    off_prep_count_ = mtake(4);
    n_fnu_ = (uint32_t)fnu.size();
    n_fmu_ = (uint32_t)fmu.size();
    off_fnu_ = mtake(fnu.size() * sizeof(PassItem));
    off_fnu_count_ = mtake(4);
    off_fmu_ = mtake(fmu.size() * sizeof(PassItem));
    off_fmu_count_ = mtake(4);
    meta_ = DeviceBuffer(std::max<size_t>(moff, kAlign));
    char* m = meta_.as<char>();
    auto put = [&](size_t off, const void* src, size_t nbytes) {
        if (nbytes) BNMF_HIP(hipMemcpy(m + off, src, nbytes, hipMemcpyHostToDevice));
    };
    put(off_ai_, ai.data(), ai.size() * sizeof(ApprInst));
    put(off_bi_, bi.data(), bi.size() * sizeof(BldInst));
    {
        size_t gi = 0;
        for (auto& kv : by_zb) {
            const Group& g = groups_[gi++];
            put(g.off_items, kv.second.data(), kv.second.size() * sizeof(PassItem));
            put(g.off_count, &g.n_items, 4);
        }
    }
    put(off_prep_, prep.data(), prep.size() * sizeof(PrepItem));
    put(off_prep_count_, &n_prep_, 4);
    put(off_fnu_, fnu.data(), fnu.size() * sizeof(PassItem));
    put(off_fnu_count_, &n_fnu_, 4);
    put(off_fmu_, fmu.data(), fmu.size() * sizeof(PassItem));
    put(off_fmu_count_, &n_fmu_, 4);
    BNMF_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
}

ApprBatch::~ApprBatch() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_) (void)hipGraphExecDestroy(graph_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void ApprBatch::enqueue_iteration() {
    char* m = meta_.as<char>();
    const ApprInst* ai = reinterpret_cast<const ApprInst*>(m + off_ai_);
    const BldInst* bi = reinterpret_cast<const BldInst*>(m + off_bi_);
    auto passes = [&](int mode) {
        for (const Group& g : groups_)
            BNMF_HIP(launch_appr_pass(dtype_ == BNMF_F32 ? 0 : 1, mode, (int)g.zb, dim3(g.n_items), stream_, ai,
                                      reinterpret_cast<const PassItem*>(m + g.off_items),
                                      reinterpret_cast<const uint32_t*>(m + g.off_count), g.zmax));
    };
    for (int half = 0; half < 2; ++half) {
        passes(0);
        BNMF_HIP(launch_prep(dtype_, psi_, false, dim3(n_prep_), stream_, bi,
                             reinterpret_cast<const PrepItem*>(m + off_prep_),
                             reinterpret_cast<const uint32_t*>(m + off_prep_count_), (uint32_t)lay_.size()));
        passes(1 + half);
        const uint32_t nf = half == 0 ? n_fnu_ : n_fmu_;
        BNMF_HIP(launch_appr_factor(dtype_ == BNMF_F32 ? 0 : 1, half, dim3(nf), stream_, ai,
                                    reinterpret_cast<const PassItem*>(m + (half == 0 ? off_fnu_ : off_fmu_)),
                                    reinterpret_cast<const uint32_t*>(m + (half == 0 ? off_fnu_count_ : off_fmu_count_))));
    }
}

void ApprBatch::capture() {
    if (graph_) return;
    hipGraph_t g = nullptr;
    BNMF_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
        enqueue_iteration();
    } catch (...) {
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    BNMF_HIP(hipStreamEndCapture(stream_, &g));
    hipError_t e = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    BNMF_HIP(e);
}

void ApprBatch::solve_async(size_t iters) {
    char* b = arena_.as<char>();
    for (const L_& l : lay_) {
        const size_t xz = (size_t)l.x * l.z, zy = (size_t)l.z * l.y;
        if (xz) BNMF_HIP(hipMemcpyAsync(b + l.off_nu, b + l.off_nu0, xz * es_, hipMemcpyDeviceToDevice, stream_));
        if (zy) BNMF_HIP(hipMemcpyAsync(b + l.off_mu, b + l.off_mu0, zy * es_, hipMemcpyDeviceToDevice, stream_));
    }
    if (n_prep_ == 0 || iters == 0) return;
    capture();
    for (size_t t = 0; t < iters; ++t) BNMF_HIP(hipGraphLaunch(graph_, stream_));
}

void ApprBatch::synchronize() { BNMF_HIP(hipStreamSynchronize(stream_)); }

void ApprBatch::get(size_t idx, void* S, void* nu, void* mu) {
    BNMF_REQUIRE(idx < lay_.size(), "instance index out of range");
    const L_& l = lay_[idx];
    char* b = arena_.as<char>();
    const size_t xy = (size_t)l.x * l.y;
    if (S && xy) {
        const ApprInst* ai = reinterpret_cast<const ApprInst*>(meta_.as<char>() + off_ai_) + idx;
        BNMF_HIP(launch_appr_S(dtype_ == BNMF_F32 ? 0 : 1, ai, xy, stream_));
    }
    BNMF_HIP(hipStreamSynchronize(stream_));
    if (S && xy) BNMF_HIP(hipMemcpy(S, b + l.off_S, xy * l.z * es_, hipMemcpyDeviceToHost));
    if (nu && l.x) BNMF_HIP(hipMemcpy(nu, b + l.off_nu, (size_t)l.x * l.z * es_, hipMemcpyDeviceToHost));
    if (mu && l.y) BNMF_HIP(hipMemcpy(mu, b + l.off_mu, (size_t)l.z * l.y * es_, hipMemcpyDeviceToHost));
}

double ApprBatch::time_iteration(int reps) {
    BNMF_REQUIRE(reps > 0, "reps must be positive");
    if (n_prep_ == 0) return 0.0;
    capture();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BNMF_HIP(hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    if (err != hipSuccess) {
        (void)hipEventDestroy(e0);
        BNMF_HIP(err);
    }
    float ms = 0.f;
    err = hipGraphLaunch(graph_, stream_);
    if (err == hipSuccess) err = hipEventRecord(e0, stream_);
    for (int r = 0; r < reps && err == hipSuccess; ++r) err = hipGraphLaunch(graph_, stream_);
    if (err == hipSuccess) err = hipEventRecord(e1, stream_);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    BNMF_HIP(err);
    return (double)ms / reps;
}

size_t ApprBatch::device_bytes() const { return arena_.bytes + meta_.bytes; }

}  // namespace bnmf
