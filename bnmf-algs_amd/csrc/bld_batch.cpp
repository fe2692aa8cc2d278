// bld_batch.cpp — planning, HBM layout and launch sequencing of the BLD
// solver (host side of the fused kernels in bld_kernels.hip).
//
// The reference driver (src/bld/bld_mult/bld_mult.hpp:226-324) allocates a
// pitched 3D buffer per tensor, copies the axis sums to the host every
// iteration and launches 7 kernels on 4 streams per iteration. Here the whole
// iteration stays on the device: one arena holds every instance's tensors,
// each iteration is `prep` + one fused `pass` per z bucket, and the chain is
// captured once into a hipGraph and replayed.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "bnmf_internal.hpp"

namespace bnmf {

namespace {
constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }
}  // namespace

// --------------------------------------------------------------- runtime
void ensure_runtime() {
    static std::once_flag once;
    static hipError_t status = hipSuccess;
    std::call_once(once, [] {
        status = init_kernel_attributes();
        if (status == hipSuccess) status = init_f32z_attributes();
        if (status == hipSuccess) status = init_f64_attributes();
        if (status == hipSuccess) status = init_appr_attributes();
    });
    BNMF_HIP(status);
}

DeviceBuffer::DeviceBuffer(size_t n) : bytes(n) {
    if (n) BNMF_HIP(hipMalloc(&ptr, n));
}
DeviceBuffer::~DeviceBuffer() {
    if (ptr) (void)hipFree(ptr);
}
DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
        if (ptr) (void)hipFree(ptr);
        ptr = o.ptr;
        bytes = o.bytes;
        o.ptr = nullptr;
        o.bytes = 0;
    }
    return *this;
}

// ---------------------------------------------------------------- planner
uint32_t z_bucket(uint32_t z) {
    if (z > kFusedZ) return 0;
    if (z <= 8) return 8;
    if (z <= 16) return 16;
    if (z <= 32) return 32;
    return 64;
}

uint32_t tile_cols_of(int dtype, uint32_t z) {
    if (z > kFusedZ) return kWideCols;
    if (dtype == BNMF_F64) return z <= 32 ? kTileColsF64 : kTileColsF64 / 2;
    return kTileCols;
}

uint32_t choose_ti_rows(int dtype, uint32_t x, uint32_t y, uint32_t z) {
    // an empty instance has no tiles (and x = 0 would divide by zero below)
    if (x == 0 || y == 0) return 4;
    // deep fibers: the wide pass's fixed tiles
    if (z > kFusedZ) return kWideRows;
    // Row-tile height depends on the instance shape only (never on the rest
    // of a batch), so an instance sums in the same order alone or batched.
    // Tall tiles (up to 128 rows) amortise the tile prologue/epilogue and
    // cut the S_+jk partial traffic to 1/TI of S; an instance with a million
    // fibers or more takes shorter tiles if that is what it needs to give
    // 256 CUs two tiles each when it runs alone (FP64, whose step is bound by
    // arithmetic rather than bytes, four).
    const uint32_t n_tj = (y + tile_cols_of(dtype, z) - 1) / tile_cols_of(dtype, z);
    // Measured on MI355X: small instances (batched) prefer 64-row tiles,
    // instances of a million fibers or more 128-row tiles. BNMF_TI_CAP
    // overrides the cap for experiments.
    static const long env_cap = [] {
        const char* e = std::getenv("BNMF_TI_CAP");
        return e ? std::strtol(e, nullptr, 10) : 0L;
    }();
    const bool large = (uint64_t)x * y >= (1ull << 20);
    // the autotuner's table entry for this shape class, if any
    const uint32_t tuned = tune_ti_cap(dtype, z, (uint64_t)x * y);
    const uint32_t cap = env_cap > 0 ? (uint32_t)std::min<long>(std::max<long>(env_cap, 4), 1024)
                                     : tuned ? tuned : (large ? 128u : 64u);
    uint32_t nt = (x + cap - 1) / cap;
    if (large) {
        const uint32_t target = dtype == BNMF_F64 ? 1024u : 512u;
        const uint32_t fill = (target + n_tj - 1) / n_tj;
        const uint32_t most = (x + 15) / 16;
        nt = std::max(nt, std::min(fill, most));
    }
    uint32_t ti = (x + nt - 1) / nt;
    ti = (ti + 3) / 4 * 4;
    return std::max<uint32_t>(ti, 4);
}

// Block b of a launch starts on XCD b % 8: the dispatcher deals workgroups
// round-robin over the eight XCDs, each with its own L2. The items of a
// launch (one block each) are laid out so that XCD c runs one contiguous
// run of the (instance, column tile, row tile) order: the row tiles of a
// column tile read the same psib slice and the column tiles of a row tile
// the same grad_minus rows, and then do so through one L2. Item order has
// no effect on results (every tile writes its own partial-sum slots).
// BNMF_TILE_ORDER=row keeps the plain row-major order.
void xcd_order(std::vector<PassItem>& v) {
    static const bool on = [] {
        const char* e = std::getenv("BNMF_TILE_ORDER");
        return !(e && std::strcmp(e, "row") == 0);
    }();
    constexpr size_t kXcds = 8;
    if (!on || v.size() < 2 * kXcds) return;
    std::stable_sort(v.begin(), v.end(), [](const PassItem& a, const PassItem& b) {
        if (a.inst != b.inst) return a.inst < b.inst;
        if (a.tj != b.tj) return a.tj < b.tj;
        return a.ti < b.ti;
    });
    const size_t n = v.size(), q = n / kXcds, r = n % kXcds;
    std::vector<PassItem> out(n);
    size_t start = 0;
    for (size_t c = 0; c < kXcds; ++c) {
        const size_t cnt = q + (c < r ? 1 : 0);
        for (size_t s = 0; s < cnt; ++s) out[s * kXcds + c] = v[start + s];
        start += cnt;
    }
    v.swap(out);
}

InstLayout plan_instance(int dtype, uint32_t x, uint32_t y, uint32_t z, bool with_S0) {
    InstLayout L{};
    const size_t es = dtype == BNMF_F32 ? 4 : 8;
    L.x = x;
    L.y = y;
    L.z = z;
    L.zb = z_bucket(z);
    read_layout(dtype, z, &L.rw, &L.hs, &L.hm);
    L.ti_rows = choose_ti_rows(dtype, x, y, z);
    L.tile_cols = tile_cols_of(dtype, z);
    L.n_ti = ceil_div(x, L.ti_rows);
    L.n_tj = ceil_div(y, L.tile_cols);
    L.nA = ceil_div(x, kPrepRows);
    L.n_cj = ceil_div(y, kPrepCols);
    size_t off = 0;
    auto take = [&](size_t n) {
        const size_t o = off;
        off = align_up(off + n);
        return o;
    };
    const size_t xyz = (size_t)x * y * z;
    // the pass reads whole 16-byte chunks around each row segment: keep
    // 64 bytes of slack after the tensors it reads
    L.off_S = take(xyz * es + 64);
    L.off_S0 = with_S0 ? take(xyz * es + 64) : L.off_S;
    L.off_X = take((size_t)x * y * es);
    L.off_alpha = take((size_t)x * 8);
    L.off_beta = take((size_t)z * 8);
    L.off_psib = take((size_t)y * z * es);
    L.off_psia = take((size_t)x * z * es);
    L.off_partA = take((size_t)L.nA * z * es);
    L.off_psiA = take((size_t)z * es);
    L.off_pjk = take((size_t)L.n_ti * y * z * 8);  // double capacity (scoring)
    L.off_ipk = take((size_t)L.n_tj * x * z * 8);
    L.off_score = take(((size_t)L.n_ti * L.n_tj + L.nA + L.n_cj + (size_t)L.nA * z) * 8);
    L.off_result = take(8);
    L.bytes = off;
    return L;
}

BldInst make_desc(const InstLayout& L, void* base, double eps, double a, double b, bool first,
                  bool nonneg) {
    char* p = static_cast<char*>(base);
    BldInst d{};
    d.S = p + L.off_S;
    d.S_src = first ? p + L.off_S0 : p + L.off_S;
    d.X = p + L.off_X;
    d.alpha = reinterpret_cast<const double*>(p + L.off_alpha);
    d.beta = reinterpret_cast<const double*>(p + L.off_beta);
    d.psib = p + L.off_psib;
    d.psia = p + L.off_psia;
    d.partA = p + L.off_partA;
    d.psiA = p + L.off_psiA;
    d.part_pjk = p + L.off_pjk;
    d.part_ipk = p + L.off_ipk;
    d.score = reinterpret_cast<double*>(p + L.off_score);
    d.result = reinterpret_cast<double*>(p + L.off_result);
    d.sijp = nullptr;
    d.eps = eps;
    d.a = a;
    d.b = b;
    d.inv_z = 1.0f / (float)L.z;
    d.x = L.x;
    d.y = L.y;
    d.z = L.z;
    d.rw = L.rw;
    d.hs = L.hs;
    d.hm = L.hm;
    d.ti_rows = L.ti_rows;
    d.tile_cols = L.tile_cols;
    d.n_ti = L.n_ti;
    d.n_tj = L.n_tj;
    d.nA = L.nA;
    d.n_cj = L.n_cj;
    uint32_t zq = 1, lg = 0;
    while (zq < L.z) {
        zq <<= 1;
        ++lg;
    }
    d.zq = zq;
    d.zq_log2 = lg;
    d.nonneg = nonneg ? 1u : 0u;
    return d;
}

PassKey pass_key(int dtype, int mode, const InstLayout& L, bool nonneg) {
    // BNMF_PASS_GENERIC=1 routes the FP32 steps through the generic z-bucket
    // kernel, for comparisons. Its update is the same formula evaluated in
    // another order (the exact-z kernel folds X_reciprocal and X / S_ij+
    // into per-fiber constants), so the two agree to FP32 rounding, not
    // bit for bit.
    static const bool generic = [] {
        const char* e = std::getenv("BNMF_PASS_GENERIC");
        return e != nullptr && e[0] == '1';
    }();
    PassKey k;
    if (L.z > kFusedZ) return k;  // the wide pass (zexact = zb = 0)
    if (dtype == BNMF_F64 && (mode == MODE_UPDATE || mode == MODE_ADD)) {
        k.f64 = 1;
        k.zb = (int)L.zb;
        k.rw = (int)L.rw;
        return k;
    }
    if (dtype == BNMF_F32 && (mode == MODE_UPDATE || mode == MODE_ADD) && nonneg && !generic) {
        k.zexact = (int)L.z;
        return k;
    }
    k.zb = (int)L.zb;
    k.rw = (mode == MODE_UPDATE || mode == MODE_ADD) ? (int)L.rw : 1;
    return k;
}

hipError_t launch_pass_key(const PassKey& k, int dtype, int psi, int mode, dim3 grid,
                           hipStream_t st, const BldInst* insts, const PassItem* items,
                           const uint32_t* n_items, uint32_t zmax) {
    if (k.wide()) return launch_pass_wide(dtype, psi, mode, grid, st, insts, items, n_items);
    if (k.f64) {
        if (dtype != BNMF_F64 || (mode != MODE_UPDATE && mode != MODE_ADD)) return hipErrorInvalidValue;
        return launch_pass_f64(psi, mode, k.zb, k.rw, grid, st, insts, items, n_items);
    }
    if (k.zexact > 0) {
        if (dtype != BNMF_F32 || (mode != MODE_UPDATE && mode != MODE_ADD)) return hipErrorInvalidValue;
        return launch_pass_f32z(k.zexact, mode, grid, st, insts, items, n_items);
    }
    return launch_pass(dtype, psi, mode, k.zb, k.rw, grid, st, insts, items, n_items, zmax);
}

// ------------------------------------------------------------------ batch
Batch::Batch(int dtype, int psi_mode, double eps, const std::vector<InstSpec>& specs, int algo)
    : dtype_(dtype), psi_(psi_mode), algo_(algo), eps_(eps), es_(dtype == BNMF_F32 ? 4 : 8), specs_(specs) {
    BNMF_REQUIRE(algo == BNMF_ALGO_MULT || algo == BNMF_ALGO_ADD, "algo must be BNMF_ALGO_MULT or BNMF_ALGO_ADD");
    // bld_add evaluates every digamma with psi_appr (src/bld/bld_add.hpp:137-146)
    if (algo == BNMF_ALGO_ADD) {
        step_mode_ = MODE_ADD;
        if (dtype == BNMF_F64) psi_ = PSIM_F64_APPR;
    }
    ensure_runtime();
    // iterations per replayed graph, up to 64 and rounded down to a power
    // of two; BNMF_GRAPH_ITERS overrides the default for experiments
    static const int env_unroll = [] {
        const char* e = std::getenv("BNMF_GRAPH_ITERS");
        const long v = e ? std::strtol(e, nullptr, 10) : 0L;
        return v > 0 ? (int)std::min<long>(v, 64) : 0;
    }();
    if (env_unroll) unroll_ = env_unroll;
    int dev = 0;
    BNMF_HIP(hipGetDevice(&dev));
    dev_ = dev;
    BNMF_HIP(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev));

    // layouts and arena
    std::vector<size_t> base_off(specs.size());
    size_t total = 0;
    for (size_t n = 0; n < specs.size(); ++n) {
        const InstSpec& s = specs[n];
        BNMF_REQUIRE(s.x < (1u << 31) && s.y < (1u << 31) && s.z < (1u << 31),
                     "x, y and z must be below 2^31");
        layouts_.push_back(plan_instance(dtype, s.x, s.y, s.z));
        base_off[n] = total;
        total += layouts_.back().bytes;
    }
    arena_ = DeviceBuffer(std::max<size_t>(total, kAlign));
    device_bytes_ = arena_.bytes;

    // uploads (X, alpha, beta, S0 or its device draw). nonneg[n]: S0 >= 0
    // and every beta_k >= 1, so grad_plus >= 0 and S stays nonnegative (the
    // kernels then skip the psi(S+1) guard for S < 0).
    std::vector<bool> nonneg(specs.size(), false);
    for (size_t n = 0; n < specs.size(); ++n) {
        const InstSpec& s = specs[n];
        const InstLayout& L = layouts_[n];
        char* b = arena_.as<char>(base_off[n]);
        const size_t xy = (size_t)s.x * s.y, xyz = L.elements();
        if (xyz == 0) continue;
        BNMF_HIP(hipMemcpy(b + L.off_X, s.X, xy * es_, hipMemcpyHostToDevice));
        BNMF_HIP(hipMemcpy(b + L.off_alpha, s.alpha, (size_t)s.x * 8, hipMemcpyHostToDevice));
        BNMF_HIP(hipMemcpy(b + L.off_beta, s.beta, (size_t)s.z * 8, hipMemcpyHostToDevice));
        const void* S0 = s.S0;
        bool nn = true;
        for (uint32_t k = 0; k < s.z; ++k) nn = nn && s.beta[k] >= 1.0;
        if (nn && s.S0) {
            if (dtype == BNMF_F32) {
                const float* p = static_cast<const float*>(S0);
                for (size_t t = 0; t < xyz && nn; ++t) nn = p[t] >= 0.0f;
            } else {
                const double* p = static_cast<const double*>(S0);
                for (size_t t = 0; t < xyz && nn; ++t) nn = p[t] >= 0.0;
            }
        }
        nonneg[n] = nn;
        if (S0) {
            BNMF_HIP(hipMemcpy(b + L.off_S0, S0, xyz * es_, hipMemcpyHostToDevice));
        } else {
            // init_S drawn on the device from the resident X (the same
            // Philox stream as the host's init_S, bit for bit): no x*y*z
            // upload
            BNMF_HIP(launch_init_s(dtype, b + L.off_S0, b + L.off_X, xy, s.z, s.seed, nullptr));
        }
        BNMF_HIP(hipMemcpy(b + L.off_S, b + L.off_S0, xyz * es_, hipMemcpyDeviceToDevice));
    }

    // descriptors and work lists of the whole batch (the sums-only and
    // scoring passes, bld_fact, tensor_sums)
    std::vector<BldInst> d_first, d_main;
    std::vector<PrepItem> prep;
    for (size_t n = 0; n < specs.size(); ++n) {
        const InstLayout& L = layouts_[n];
        char* b = arena_.as<char>(base_off[n]);
        d_first.push_back(make_desc(L, b, eps, specs[n].a, specs[n].b, true, nonneg[n]));
        d_main.push_back(make_desc(L, b, eps, specs[n].a, specs[n].b, false, nonneg[n]));
        if (L.elements() == 0) continue;
        for (uint32_t c = 0; c < L.n_cj; ++c) prep.push_back(PrepItem{(uint32_t)n, 0, c, 0});
        for (uint32_t c = 0; c < L.nA; ++c) prep.push_back(PrepItem{(uint32_t)n, 1, c, 0});
    }
    // one pass launch per kernel key, separately for the update pass and for
    // the sums-only / scoring passes
    std::map<PassKey, std::vector<PassItem>> sum_items;
    std::map<PassKey, std::vector<uint32_t>> members;  // update groups: batch instance numbers
    for (size_t n = 0; n < specs.size(); ++n) {
        const InstLayout& L = layouts_[n];
        if (L.elements() == 0) continue;
        auto& v = sum_items[pass_key(dtype, MODE_SUMS, L, nonneg[n])];
        for (uint32_t ti = 0; ti < L.n_ti; ++ti)
            for (uint32_t tj = 0; tj < L.n_tj; ++tj) v.push_back(PassItem{(uint32_t)n, ti, tj, 0});
        members[pass_key(dtype, step_mode_, L, nonneg[n])].push_back((uint32_t)n);
    }
    auto tile_work = [&](const InstLayout& L, uint32_t ti, uint32_t tj) {
        const uint32_t rows = std::min(L.ti_rows, L.x - ti * L.ti_rows);
        const uint32_t cols = std::min(L.tile_cols, L.y - tj * L.tile_cols);
        return (size_t)rows * cols * L.z;
    };
    for (auto& kv : sum_items) {
        xcd_order(kv.second);
        Bucket bk;
        bk.key = kv.first;
        bk.zmax = 1;
        bk.work = 0;
        for (const PassItem& pi : kv.second) {
            bk.work += tile_work(layouts_[pi.inst], pi.ti, pi.tj);
            bk.zmax = std::max(bk.zmax, layouts_[pi.inst].z);
        }
        bk.lane = 0;
        bk.n_items = (uint32_t)kv.second.size();
        groups_[1].push_back(bk);
    }
    for (auto& kv : members) {
        Bucket bk;
        bk.key = kv.first;
        bk.zmax = 1;
        bk.work = 0;
        for (uint32_t n : kv.second) {
            bk.work += layouts_[n].elements();
            bk.zmax = std::max(bk.zmax, layouts_[n].z);
        }
        bk.lane = 0;
        groups_[0].push_back(bk);
    }
    // up to four streams (BNMF_LANES: 1 to 16); each group goes to the least
    // loaded one, largest work first
    static const size_t max_lanes = [] {
        const char* e = std::getenv("BNMF_LANES");
        const long v = e ? std::strtol(e, nullptr, 10) : 0L;
        return v > 0 ? (size_t)std::min<long>(v, 16) : (size_t)4;
    }();
    const int ns = (int)std::min<size_t>(max_lanes, std::max(groups_[0].size(), groups_[1].size()));
    if (ns > 1) {
        for (int g = 0; g < 2; ++g) {
            std::vector<Bucket>& gs = groups_[g];
            std::vector<size_t> order(gs.size()), load(ns, 0);
            for (size_t b = 0; b < order.size(); ++b) order[b] = b;
            std::stable_sort(order.begin(), order.end(),
                             [&](size_t a, size_t b) { return gs[a].work > gs[b].work; });
            for (size_t b : order) {
                const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                gs[b].lane = l;
                load[l] += gs[b].work;
            }
        }
    }
    // Each stream (lane) has its own descriptor array and prep work list
    // over the instances of its update groups; their pass items use the
    // lane's instance numbers.
    lanes_.assign(std::max(ns, 1), Lane{});
    std::vector<std::vector<uint32_t>> lane_members(lanes_.size());
    std::vector<std::vector<PassItem>> upd_items(groups_[0].size());
    {
        size_t gi = 0;
        for (auto& kv : members) {
            Bucket& bk = groups_[0][gi];
            std::vector<uint32_t>& lm = lane_members[bk.lane];
            for (uint32_t n : kv.second) {
                const uint32_t local = (uint32_t)lm.size();
                lm.push_back(n);
                const InstLayout& L = layouts_[n];
                for (uint32_t ti = 0; ti < L.n_ti; ++ti)
                    for (uint32_t tj = 0; tj < L.n_tj; ++tj) upd_items[gi].push_back(PassItem{local, ti, tj, 0});
            }
            bk.n_items = (uint32_t)upd_items[gi].size();
            xcd_order(upd_items[gi]);
            ++gi;
        }
    }
    size_t moff = 0;
    auto mtake = [&](size_t nbytes) {
        const size_t o = moff;
        moff = align_up(moff + nbytes);
        return o;
    };
    off_desc_first_ = mtake(d_first.size() * sizeof(BldInst));
    off_desc_ = mtake(d_main.size() * sizeof(BldInst));
    off_iter_ = mtake(specs.size() * sizeof(uint32_t));
    off_prep_items_ = mtake(prep.size() * sizeof(PrepItem));
    off_prep_count_ = mtake(sizeof(uint32_t));
    n_prep_items_ = (uint32_t)prep.size();
    for (int g = 0; g < 2; ++g)
        for (Bucket& bk : groups_[g]) {
            bk.off_items = mtake((size_t)bk.n_items * sizeof(PassItem));
            bk.off_count = mtake(sizeof(uint32_t));
        }
    for (size_t l = 0; l < lanes_.size(); ++l) {
        Lane& ln = lanes_[l];
        ln.n_inst = (uint32_t)lane_members[l].size();
        for (uint32_t n : lane_members[l]) ln.n_prep += layouts_[n].n_cj + layouts_[n].nA;
        ln.off_desc = mtake(ln.n_inst * sizeof(BldInst));
        ln.off_desc_first = mtake(ln.n_inst * sizeof(BldInst));
        ln.off_prep = mtake((size_t)ln.n_prep * sizeof(PrepItem));
        ln.off_prep_count = mtake(sizeof(uint32_t));
    }
    meta_ = DeviceBuffer(std::max<size_t>(moff, kAlign));
    device_bytes_ += meta_.bytes;
    char* m = meta_.as<char>();
    if (algo_ == BNMF_ALGO_ADD) {
        for (size_t n = 0; n < specs.size(); ++n) {
            uint32_t* ctr = reinterpret_cast<uint32_t*>(m + off_iter_) + n;
            d_first[n].iter = ctr;
            d_main[n].iter = ctr;
        }
        BNMF_HIP(hipMemset(m + off_iter_, 0, specs.size() * sizeof(uint32_t)));
    }
    if (!d_first.empty()) {
        BNMF_HIP(hipMemcpy(m + off_desc_first_, d_first.data(), d_first.size() * sizeof(BldInst), hipMemcpyHostToDevice));
        BNMF_HIP(hipMemcpy(m + off_desc_, d_main.data(), d_main.size() * sizeof(BldInst), hipMemcpyHostToDevice));
    }
    if (!prep.empty())
        BNMF_HIP(hipMemcpy(m + off_prep_items_, prep.data(), prep.size() * sizeof(PrepItem), hipMemcpyHostToDevice));
    BNMF_HIP(hipMemcpy(m + off_prep_count_, &n_prep_items_, sizeof(uint32_t), hipMemcpyHostToDevice));
    {
        size_t gi = 0;
        for (auto& kv : sum_items) {
            Bucket& bk = groups_[1][gi++];
            BNMF_HIP(hipMemcpy(m + bk.off_items, kv.second.data(), kv.second.size() * sizeof(PassItem), hipMemcpyHostToDevice));
            BNMF_HIP(hipMemcpy(m + bk.off_count, &bk.n_items, sizeof(uint32_t), hipMemcpyHostToDevice));
        }
        for (gi = 0; gi < groups_[0].size(); ++gi) {
            Bucket& bk = groups_[0][gi];
            BNMF_HIP(hipMemcpy(m + bk.off_items, upd_items[gi].data(), upd_items[gi].size() * sizeof(PassItem), hipMemcpyHostToDevice));
            BNMF_HIP(hipMemcpy(m + bk.off_count, &bk.n_items, sizeof(uint32_t), hipMemcpyHostToDevice));
        }
    }
    for (size_t l = 0; l < lanes_.size(); ++l) {
        const Lane& ln = lanes_[l];
        std::vector<BldInst> dm, df;
        std::vector<PrepItem> lp;
        for (uint32_t loc = 0; loc < ln.n_inst; ++loc) {
            const uint32_t n = lane_members[l][loc];
            dm.push_back(d_main[n]);
            df.push_back(d_first[n]);
            for (uint32_t c = 0; c < layouts_[n].n_cj; ++c) lp.push_back(PrepItem{loc, 0, c, 0});
            for (uint32_t c = 0; c < layouts_[n].nA; ++c) lp.push_back(PrepItem{loc, 1, c, 0});
        }
        if (ln.n_inst) {
            BNMF_HIP(hipMemcpy(m + ln.off_desc, dm.data(), dm.size() * sizeof(BldInst), hipMemcpyHostToDevice));
            BNMF_HIP(hipMemcpy(m + ln.off_desc_first, df.data(), df.size() * sizeof(BldInst), hipMemcpyHostToDevice));
        }
        if (ln.n_prep)
            BNMF_HIP(hipMemcpy(m + ln.off_prep, lp.data(), lp.size() * sizeof(PrepItem), hipMemcpyHostToDevice));
        BNMF_HIP(hipMemcpy(m + ln.off_prep_count, &ln.n_prep, sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    BNMF_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    if (ns > 1) {
        side_.resize(ns);
        fork_ev_.resize(ns);
        join_ev_.resize(ns);
        side_[0] = stream_;
        for (int l = 1; l < ns; ++l) BNMF_HIP(hipStreamCreateWithFlags(&side_[l], hipStreamNonBlocking));
        for (int l = 0; l < ns; ++l) {
            BNMF_HIP(hipEventCreateWithFlags(&fork_ev_[l], hipEventDisableTiming));
            BNMF_HIP(hipEventCreateWithFlags(&join_ev_[l], hipEventDisableTiming));
        }
    }
}

Batch::~Batch() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (hipGraphExec_t g : graphs_)
        if (g) (void)hipGraphExecDestroy(g);
    for (size_t l = 1; l < side_.size(); ++l) (void)hipStreamDestroy(side_[l]);
    for (hipEvent_t e : fork_ev_) (void)hipEventDestroy(e);
    for (hipEvent_t e : join_ev_) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Batch::enqueue_pass(int mode, bool first) {
    char* m = meta_.as<char>();
    const BldInst* d = reinterpret_cast<const BldInst*>(m + (first ? off_desc_first_ : off_desc_));
    const bool fork = side_.size() > 1;
    if (fork) {
        BNMF_HIP(hipEventRecord(fork_ev_[0], stream_));
        for (size_t l = 1; l < side_.size(); ++l) BNMF_HIP(hipStreamWaitEvent(side_[l], fork_ev_[0], 0));
    }
    for (const Bucket& bk : groups_of(mode)) {
        const dim3 grid(bk.n_items);
        hipStream_t st = fork ? side_[bk.lane] : stream_;
        const BldInst* dg =
            (mode == MODE_UPDATE || mode == MODE_ADD)
                ? reinterpret_cast<const BldInst*>(
                      m + (first ? lanes_[bk.lane].off_desc_first : lanes_[bk.lane].off_desc))
                : d;
        BNMF_HIP(launch_pass_key(bk.key, dtype_, psi_, mode, grid, st, dg,
                                 reinterpret_cast<const PassItem*>(m + bk.off_items),
                                 reinterpret_cast<const uint32_t*>(m + bk.off_count), bk.zmax));
    }
    if (fork) {
        for (size_t l = 1; l < side_.size(); ++l) {
            BNMF_HIP(hipEventRecord(join_ev_[l], side_[l]));
            BNMF_HIP(hipStreamWaitEvent(stream_, join_ev_[l], 0));
        }
    }
}

void Batch::enqueue_iterations(int iters, bool first, bool with_prep) {
    if (n_prep_items_ == 0 || iters <= 0) return;
    char* m = meta_.as<char>();
    const bool fork = side_.size() > 1;
    if (fork) {
        BNMF_HIP(hipEventRecord(fork_ev_[0], stream_));
        for (size_t l = 1; l < side_.size(); ++l) BNMF_HIP(hipStreamWaitEvent(side_[l], fork_ev_[0], 0));
    }
    // Instances of different streams never interact, so each stream runs
    // its own iterations back to back (prep and psi(A) of its instances,
    // then its groups' passes): one stream's prep overlaps the others'
    // passes and no stream waits for the others between iterations.
    for (int t = 0; t < iters; ++t) {
        for (size_t l = 0; l < lanes_.size(); ++l) {
            const Lane& ln = lanes_[l];
            if (ln.n_inst == 0) continue;
            hipStream_t st = fork ? side_[l] : stream_;
            if (with_prep)
                BNMF_HIP(launch_prep(dtype_, psi_, false, dim3(ln.n_prep), st,
                                     reinterpret_cast<const BldInst*>(m + ln.off_desc),
                                     reinterpret_cast<const PrepItem*>(m + ln.off_prep),
                                     reinterpret_cast<const uint32_t*>(m + ln.off_prep_count), ln.n_inst));
            const BldInst* dp =
                reinterpret_cast<const BldInst*>(m + ((first && t == 0) ? ln.off_desc_first : ln.off_desc));
            for (const Bucket& bk : groups_[0]) {
                if ((size_t)bk.lane != l) continue;
                BNMF_HIP(launch_pass_key(bk.key, dtype_, psi_, step_mode_, dim3(bk.n_items), st, dp,
                                         reinterpret_cast<const PassItem*>(m + bk.off_items),
                                         reinterpret_cast<const uint32_t*>(m + bk.off_count), bk.zmax));
            }
        }
    }
    if (fork) {
        for (size_t l = 1; l < side_.size(); ++l) {
            BNMF_HIP(hipEventRecord(join_ev_[l], side_[l]));
            BNMF_HIP(hipStreamWaitEvent(stream_, join_ev_[l], 0));
        }
    }
}

hipGraphExec_t Batch::graph_of(int level) {
    if (graphs_[level]) return graphs_[level];
    hipGraph_t g = nullptr;
    BNMF_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
        enqueue_iterations(1 << level, false);
    } catch (...) {
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    BNMF_HIP(hipStreamEndCapture(stream_, &g));
    hipError_t e = hipGraphInstantiate(&graphs_[level], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    BNMF_HIP(e);
    return graphs_[level];
}

void Batch::solve_async(size_t iters) {
    partials_valid_ = false;
    if (n_prep_items_ == 0) return;
    if (iters == 0) {
        // S = S0 for every instance
        size_t base = 0;
        for (const InstLayout& L : layouts_) {
            char* b = arena_.as<char>(base);
            BNMF_HIP(hipMemcpyAsync(b + L.off_S, b + L.off_S0, L.elements() * es_,
                                    hipMemcpyDeviceToDevice, stream_));
            base += L.bytes;
        }
        return;
    }
    start_solve();
    replay_iterations(iters - 1);
    partials_valid_ = true;
}

void Batch::start_solve() {
    partials_valid_ = false;
    // the additive step size restarts with the solve
    if (algo_ == BNMF_ALGO_ADD)
        BNMF_HIP(hipMemsetAsync(meta_.as<char>() + off_iter_, 0, specs_.size() * sizeof(uint32_t), stream_));
    // partial axis sums of S0, then the first iteration reading S0
    enqueue_pass(MODE_SUMS, true);
    enqueue_iteration(true);
}

void Batch::replay_iterations(size_t rem) {
    // whole graphs of unroll_ iterations, then the rest in binary: every
    // graph boundary joins the streams, so a solve has as few as it can
    int top = 0;
    while ((2 << top) <= unroll_ && top + 1 < kGraphLevels) ++top;
    const size_t big = (size_t)1 << top;
    while (rem >= big) {
        BNMF_HIP(hipGraphLaunch(graph_of(top), stream_));
        rem -= big;
    }
    for (int l = top - 1; l >= 0; --l)
        if (rem & ((size_t)1 << l)) BNMF_HIP(hipGraphLaunch(graph_of(l), stream_));
}

size_t Batch::solve_tol(size_t max_iter, double tol, size_t check_every, double* lm_out) {
    BNMF_REQUIRE(check_every >= 1, "check_every must be at least 1");
    BNMF_REQUIRE(tol >= 0.0, "tol must be nonnegative");
    const size_t n = layouts_.size();
    if (n_prep_items_ == 0 || max_iter == 0) {
        solve_async(max_iter);
        synchronize();
        if (lm_out) log_marginal(lm_out);
        return max_iter;
    }
    std::vector<double> prev(n, std::nan("")), cur(n);
    size_t done = 0;
    bool scored = false;
    while (done < max_iter) {
        const size_t chunk = std::min(check_every, max_iter - done);
        if (done == 0) {
            start_solve();
            replay_iterations(chunk - 1);
        } else {
            // the scoring pass reused the partial buffers: the partials of
            // the current S first (the same tiles and order as the step
            // pass's, so the continued solve is the uninterrupted one)
            enqueue_pass(MODE_SUMS, false);
            replay_iterations(chunk);
        }
        done += chunk;
        scored = false;
        if (done >= max_iter) break;
        log_marginal(cur.data());
        scored = true;
        bool conv = true;
        for (size_t t = 0; t < n; ++t)
            conv = conv && std::fabs(cur[t] - prev[t]) <= tol * std::fabs(cur[t]);
        prev = cur;
        if (conv) break;
    }
    if (!scored) {
        synchronize();
        partials_valid_ = true;
        if (lm_out) log_marginal(lm_out);
    } else if (lm_out) {
        std::copy(cur.begin(), cur.end(), lm_out);
    }
    return done;
}

void Batch::synchronize() { BNMF_HIP(hipStreamSynchronize(stream_)); }

void Batch::get_S(size_t idx, void* host) const {
    BNMF_REQUIRE(idx < layouts_.size(), "instance index out of range");
    size_t base = 0;
    for (size_t n = 0; n < idx; ++n) base += layouts_[n].bytes;
    const InstLayout& L = layouts_[idx];
    if (L.elements() == 0) return;
    BNMF_HIP(hipStreamSynchronize(stream_));
    BNMF_HIP(hipMemcpy(host, arena_.as<char>(base) + L.off_S, L.elements() * es_, hipMemcpyDeviceToHost));
}

void Batch::log_marginal(double* out) {
    partials_valid_ = false;
    if (n_prep_items_ > 0) {
        enqueue_pass(MODE_SCORE, false);
        char* m = meta_.as<char>();
        BNMF_HIP(launch_score(dim3(n_prep_items_), stream_, reinterpret_cast<const BldInst*>(m + off_desc_),
                              reinterpret_cast<const PrepItem*>(m + off_prep_items_),
                              reinterpret_cast<const uint32_t*>(m + off_prep_count_),
                              (uint32_t)layouts_.size()));
    }
    BNMF_HIP(hipStreamSynchronize(stream_));
    size_t base = 0;
    for (size_t n = 0; n < layouts_.size(); ++n) {
        const InstLayout& L = layouts_[n];
        if (L.elements() == 0) {
            out[n] = 0.0;
        } else {
            BNMF_HIP(hipMemcpy(&out[n], arena_.as<char>(base) + L.off_result, 8, hipMemcpyDeviceToHost));
        }
        base += L.bytes;
    }
}

void Batch::time_pass(int reps, double* pass_ms, double* iter_ms) {
    BNMF_REQUIRE(reps > 0, "reps must be positive");
    if (n_prep_items_ == 0) {
        *pass_ms = *iter_ms = 0.0;
        return;
    }
    partials_valid_ = false;
    // Two graphs, launched the way solves launch them: reps update passes
    // (every pass launch of an iteration, each group's on its stream), and
    // reps whole iterations (prep + pass). Each is replayed once to warm up,
    // then timed with events on the batch stream.
    auto capture = [&](bool with_prep, hipGraphExec_t* out) {
        hipGraph_t g = nullptr;
        BNMF_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_iterations(reps, false, with_prep);
        } catch (...) {
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        BNMF_HIP(hipStreamEndCapture(stream_, &g));
        hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        BNMF_HIP(e);
    };
    hipGraphExec_t gp = nullptr, gi = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    auto cleanup = [&] {
        if (gp) (void)hipGraphExecDestroy(gp);
        if (gi) (void)hipGraphExecDestroy(gi);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (e2) (void)hipEventDestroy(e2);
    };
    try {
        capture(false, &gp);
        capture(true, &gi);
        BNMF_HIP(hipEventCreate(&e0));
        BNMF_HIP(hipEventCreate(&e1));
        BNMF_HIP(hipEventCreate(&e2));
        BNMF_HIP(hipGraphLaunch(gp, stream_));
        BNMF_HIP(hipGraphLaunch(gi, stream_));
        BNMF_HIP(hipEventRecord(e0, stream_));
        BNMF_HIP(hipGraphLaunch(gp, stream_));
        BNMF_HIP(hipEventRecord(e1, stream_));
        BNMF_HIP(hipGraphLaunch(gi, stream_));
        BNMF_HIP(hipEventRecord(e2, stream_));
        BNMF_HIP(hipEventSynchronize(e2));
        float tp = 0.f, ti = 0.f;
        BNMF_HIP(hipEventElapsedTime(&tp, e0, e1));
        BNMF_HIP(hipEventElapsedTime(&ti, e1, e2));
        *pass_ms = (double)tp / reps;
        *iter_ms = (double)ti / reps;
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

void Batch::tensor_sums(void* S_pjk, void* S_ipk, void* S_ijp) {
    BNMF_REQUIRE(layouts_.size() == 1, "tensor_sums works on a single instance");
    partials_valid_ = false;
    const InstLayout& L = layouts_[0];
    if (L.elements() == 0) return;
    // sums of S (the current tensor), fiber sums into a scratch buffer
    DeviceBuffer sijp((size_t)L.x * L.y * es_);
    BldInst d = make_desc(L, arena_.ptr, eps_, 1.0, 1.0, false);
    d.sijp = sijp.ptr;
    DeviceBuffer dd(sizeof(BldInst));
    BNMF_HIP(hipMemcpy(dd.ptr, &d, sizeof(BldInst), hipMemcpyHostToDevice));
    char* m = meta_.as<char>();
    const Bucket& bk = groups_of(MODE_SUMS).at(0);
    BNMF_HIP(launch_pass_key(bk.key, dtype_, psi_, MODE_SUMS, dim3(bk.n_items), stream_,
                             dd.as<BldInst>(), reinterpret_cast<const PassItem*>(m + bk.off_items),
                             reinterpret_cast<const uint32_t*>(m + bk.off_count), bk.zmax));
    BNMF_HIP(launch_prep(dtype_, psi_, true, dim3(n_prep_items_), stream_, dd.as<BldInst>(),
                         reinterpret_cast<const PrepItem*>(m + off_prep_items_),
                         reinterpret_cast<const uint32_t*>(m + off_prep_count_), (uint32_t)layouts_.size()));
    BNMF_HIP(hipStreamSynchronize(stream_));
    if (S_pjk) BNMF_HIP(hipMemcpy(S_pjk, arena_.as<char>(L.off_psib), (size_t)L.y * L.z * es_, hipMemcpyDeviceToHost));
    if (S_ipk) BNMF_HIP(hipMemcpy(S_ipk, arena_.as<char>(L.off_psia), (size_t)L.x * L.z * es_, hipMemcpyDeviceToHost));
    if (S_ijp) BNMF_HIP(hipMemcpy(S_ijp, sijp.ptr, (size_t)L.x * L.y * es_, hipMemcpyDeviceToHost));
}

void Batch::fact(size_t idx, double eps, double* W, double* H, double* L) {
    BNMF_REQUIRE(idx < layouts_.size(), "instance index out of range");
    const InstLayout& Li = layouts_[idx];
    if (Li.elements() == 0) return;
    char* m = meta_.as<char>();
    if (!partials_valid_) {
        // the partials of the current S (one read of S)
        enqueue_pass(MODE_SUMS, false);
        partials_valid_ = true;
    }
    // raw prep: S_+jk and S_i+k into psib / psia (the next solve starts
    // from S0 and recomputes them); the partials stay as they were
    BNMF_HIP(launch_prep(dtype_, psi_, true, dim3(n_prep_items_), stream_,
                         reinterpret_cast<const BldInst*>(m + off_desc_),
                         reinterpret_cast<const PrepItem*>(m + off_prep_items_),
                         reinterpret_cast<const uint32_t*>(m + off_prep_count_), (uint32_t)layouts_.size()));
    const size_t nW = (size_t)Li.x * Li.z, nH = (size_t)Li.z * Li.y, nL = Li.y;
    DeviceBuffer out((nW + nH + nL) * 8);
    const BldInst* d = reinterpret_cast<const BldInst*>(m + off_desc_) + idx;
    BNMF_HIP(launch_fact(dtype_, d, Li.y, Li.z, eps, out.as<double>(), out.as<double>(nW * 8),
                         out.as<double>((nW + nH) * 8), stream_));
    BNMF_HIP(hipStreamSynchronize(stream_));
    if (W) BNMF_HIP(hipMemcpy(W, out.as<double>(), nW * 8, hipMemcpyDeviceToHost));
    if (H) BNMF_HIP(hipMemcpy(H, out.as<double>(nW * 8), nH * 8, hipMemcpyDeviceToHost));
    if (L) BNMF_HIP(hipMemcpy(L, out.as<double>((nW + nH) * 8), nL * 8, hipMemcpyDeviceToHost));
}

bnmf_batch_info Batch::info() const {
    bnmf_batch_info in{};
    in.n_instances = layouts_.size();
    for (const InstLayout& L : layouts_) {
        in.elements += L.elements();
        in.fibers += (size_t)L.x * L.y;
    }
    in.device_bytes = device_bytes_;
    for (const Bucket& bk : groups_[0]) in.pass_items += bk.n_items;
    in.prep_items = n_prep_items_;
    in.n_buckets = (int)groups_[0].size();
    return in;
}

}  // namespace bnmf
