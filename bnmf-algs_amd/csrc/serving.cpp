// serving.cpp — continuous-batching solve server for bld_mult requests.
//
// No counterpart in the reference (bld_mult solves one instance per call,
// src/bld/bld_mult/bld_mult.hpp:83-150). A request is one bld_mult problem
// (X, z, params, max_iter). The server:
//  * owns one HBM arena sized by the caller's budget; a request is admitted
//    when its tensors (S plus the per-instance workspace: psi tables and the
//    partial axis sums) fit in the free space (first fit), FIFO order;
//  * keeps S resident in HBM from admission until the client collects it;
//  * advances every active request together: each step runs a few
//    iterations of prep + fused pass over the union of the active
//    instances' tiles; requests that arrive meanwhile join at the next step
//    boundary, finished ones leave;
//  * draws S0 on the device from the same Philox stream as the host init_S
//    when the request carries none, so only X crosses PCIe.
// Every instance's tiles and sums are fixed by its own shape, so a result is
// bitwise the same whichever requests share the GPU with it.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "arena.hpp"
#include "bnmf_internal.hpp"

namespace bnmf {
namespace {

constexpr size_t kAlign = Arena::kAlign;
size_t align_up(size_t v) { return Arena::align_up(v); }

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

enum ReqState { RS_QUEUED = 0, RS_ACTIVE = 1, RS_DONE = 2, RS_FAILED = 3 };

struct Request {
    uint64_t ticket = 0;
    uint32_t x = 0, y = 0, z = 0;
    std::vector<char> X, S0;
    std::vector<double> alpha, beta;
    double a = 1, b = 1;
    uint64_t seed = 0;
    bool has_S0 = false, want_lm = false;
    bool nonneg = false;  // S0 >= 0 and beta >= 1: S stays >= 0
    size_t max_iter = 0, done_iter = 0;
    InstLayout L{};
    size_t off = 0;
    int state = RS_QUEUED;
    bool fresh = true;  // needs the partial sums of S0 before its first iteration
    double log_marginal = NAN;
    double t_submit = 0, t_admit = 0, t_done = 0;
    bnmf_status status = BNMF_OK;
    std::string err;
};

}  // namespace

class Server {
   public:
    Server(int dtype, size_t budget, size_t max_active, int psi, double eps)
        : dtype_(dtype), psi_(psi), eps_(eps), es_(dtype == BNMF_F32 ? 4 : 8),
          max_active_(max_active ? max_active : 4096), arena_(align_up(budget)) {
        ensure_runtime();
        BNMF_REQUIRE(budget >= kAlign, "HBM budget is too small");
        dev_buf_ = DeviceBuffer(arena_.capacity());
        BNMF_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        side_.resize(4);
        side_[0] = stream_;
        for (int l = 1; l < 4; ++l) BNMF_HIP(hipStreamCreateWithFlags(&side_[l], hipStreamNonBlocking));
        fork_.resize(4);
        join_.resize(4);
        for (int l = 0; l < 4; ++l) {
            BNMF_HIP(hipEventCreateWithFlags(&fork_[l], hipEventDisableTiming));
            BNMF_HIP(hipEventCreateWithFlags(&join_[l], hipEventDisableTiming));
        }
        int dev = 0;
        BNMF_HIP(hipGetDevice(&dev));
        dev_ = dev;
        th_ = std::thread([this] { loop(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
        if (stream_) (void)hipStreamSynchronize(stream_);
        for (size_t l = 1; l < side_.size(); ++l) (void)hipStreamDestroy(side_[l]);
        for (hipEvent_t e : fork_) (void)hipEventDestroy(e);
        for (hipEvent_t e : join_) (void)hipEventDestroy(e);
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    static size_t request_bytes(int dtype, uint32_t x, uint32_t y, uint32_t z) {
        return plan_instance(dtype, x, y, z, false).bytes;
    }

    uint64_t submit(const bnmf_bld_desc& d, size_t max_iter, int want_lm) {
        // check_bld_params (src/bld/util_details.hpp:20-28): X >= 0
        const size_t nxy = d.x * d.y;
        for (size_t t = 0; t < nxy; ++t) {
            const double v = dtype_ == BNMF_F32 ? (double)static_cast<const float*>(d.X)[t]
                                                : static_cast<const double*>(d.X)[t];
            BNMF_REQUIRE(v >= 0.0, "X must be nonnegative");
        }
        auto r = std::make_shared<Request>();
        r->x = (uint32_t)d.x;
        r->y = (uint32_t)d.y;
        r->z = (uint32_t)d.z;
        const size_t xy = d.x * d.y, xyz = xy * d.z;
        r->X.assign(static_cast<const char*>(d.X), static_cast<const char*>(d.X) + xy * es_);
        r->alpha.assign(d.alpha, d.alpha + d.x);
        r->beta.assign(d.beta, d.beta + d.z);
        r->a = d.a;
        r->b = d.b;
        r->seed = d.seed;
        bool nn = true;
        for (size_t k = 0; k < d.z; ++k) nn = nn && d.beta[k] >= 1.0;
        if (d.S0) {
            r->has_S0 = true;
            r->S0.assign(static_cast<const char*>(d.S0), static_cast<const char*>(d.S0) + xyz * es_);
            for (size_t t = 0; t < xyz && nn; ++t)
                nn = dtype_ == BNMF_F32 ? static_cast<const float*>(d.S0)[t] >= 0.0f
                                        : static_cast<const double*>(d.S0)[t] >= 0.0;
        }
        r->nonneg = nn;
        r->want_lm = want_lm != 0;
        r->max_iter = max_iter;
        r->L = plan_instance(dtype_, r->x, r->y, r->z, false);
        r->t_submit = now_s();
        std::lock_guard<std::mutex> lk(mu_);
        if (failed_) throw Error(BNMF_ERR_DEVICE, "server failed: " + fail_msg_);
        r->ticket = ++next_ticket_;
        ++submitted_;
        if (r->L.bytes > arena_.capacity()) {
            r->state = RS_FAILED;
            r->status = BNMF_ERR_REJECTED;
            r->err = "request needs " + std::to_string(r->L.bytes) + " bytes of HBM, above the budget";
            r->t_done = r->t_submit;
            ++rejected_;
        } else {
            queue_.push_back(r);
        }
        all_[r->ticket] = r;
        cv_.notify_all();
        return r->ticket;
    }

    bool poll(uint64_t t, bnmf_status* st) {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = all_.find(t);
        if (it == all_.end()) throw Error(BNMF_ERR_NOT_FOUND, "unknown ticket");
        *st = it->second->status;
        return it->second->state >= RS_DONE;
    }

    bnmf_status wait(uint64_t t, double timeout_s) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = all_.find(t);
        if (it == all_.end()) throw Error(BNMF_ERR_NOT_FOUND, "unknown ticket");
        auto r = it->second;
        auto pred = [&] { return r->state >= RS_DONE; };
        if (timeout_s < 0) {
            done_cv_.wait(lk, pred);
        } else if (!done_cv_.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) {
            return BNMF_ERR_TIMEOUT;
        }
        return r->status;
    }

    void collect(uint64_t t, void* S_out, double* lm, bnmf_request_stats* st) {
        std::shared_ptr<Request> r;
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = all_.find(t);
            if (it == all_.end()) throw Error(BNMF_ERR_NOT_FOUND, "unknown ticket");
            r = it->second;
            if (r->state < RS_DONE) throw Error(BNMF_ERR_NOT_READY, "request not finished");
        }
        if (r->state == RS_DONE && S_out && r->L.elements()) {
            BNMF_HIP(hipSetDevice(dev_));
            BNMF_HIP(hipMemcpy(S_out, dev_buf_.as<char>(r->off) + r->L.off_S, r->L.elements() * es_,
                               hipMemcpyDeviceToHost));
        }
        if (lm) *lm = r->log_marginal;
        if (st) {
            st->submit_s = r->t_submit;
            st->admit_s = r->t_admit;
            st->done_s = r->t_done;
            st->iterations = r->done_iter;
            st->device_bytes = r->L.bytes;
        }
        std::lock_guard<std::mutex> lk(mu_);
        if (r->state == RS_DONE) arena_.release(r->off, r->L.bytes);
        all_.erase(t);
        const bnmf_status s = r->status;
        cv_.notify_all();  // memory freed: the scheduler may admit
        if (s != BNMF_OK) throw Error(s, r->err);
    }

    bnmf_server_stats stats() {
        std::lock_guard<std::mutex> lk(mu_);
        bnmf_server_stats s{};
        s.submitted = submitted_;
        s.completed = completed_;
        s.rejected = rejected_;
        s.active = active_.size();
        s.queued = queue_.size();
        s.steps = steps_;
        s.bytes_in_use = arena_.used();
        s.peak_bytes = peak_bytes_;
        s.peak_active = peak_active_;
        s.busy_s = busy_s_;
        return s;
    }

   private:
    struct Launch {
        PassKey key;
        int lane;
        uint32_t zmax, n;
        size_t off_items, off_count, work;
    };
    // Device-side work lists for a set of instances.
    // Update work sets also keep, per stream, a descriptor array and prep
    // list over the instances of that stream's launches (their pass items
    // index it), so each stream can iterate on its own.
    struct LaneWS {
        size_t off_desc = 0, off_prep = 0, off_prep_count = 0;
        uint32_t n_prep = 0, n_inst = 0;
    };
    struct WorkSet {
        std::vector<Launch> launches;
        size_t off_desc = 0, off_prep = 0, off_prep_count = 0;
        uint32_t n_prep = 0, n_inst = 0;
        LaneWS lanes[4];
    };

    void loop() {
        (void)hipSetDevice(dev_);
        std::vector<std::shared_ptr<Request>> admitted;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !active_.empty() || can_admit_locked(); });
                if (stop_) return;
                admitted.clear();
                while (!queue_.empty() && active_.size() + admitted.size() < max_active_) {
                    auto& r = queue_.front();
                    size_t off;
                    if (!arena_.alloc(r->L.bytes, &off)) break;
                    r->off = off;
                    r->state = RS_ACTIVE;
                    r->t_admit = now_s();
                    admitted.push_back(r);
                    queue_.pop_front();
                }
                peak_bytes_ = std::max(peak_bytes_, arena_.used());
            }
            try {
                for (auto& r : admitted) upload(*r);
                std::vector<std::shared_ptr<Request>> act;
                {
                    std::lock_guard<std::mutex> lk(mu_);
                    for (auto& r : admitted) active_.push_back(r);
                    act = active_;
                    peak_active_ = std::max(peak_active_, act.size());
                }
                if (act.empty()) continue;
                step(act);
            } catch (const std::exception& e) {
                fail(e.what());
                return;
            }
        }
    }

    // The queue head can start now: a slot is free and its bytes fit.
    bool can_admit_locked() const {
        return !queue_.empty() && active_.size() < max_active_ && arena_.fits(queue_.front()->L.bytes);
    }

    void upload(Request& r) {
        char* base = dev_buf_.as<char>(r.off);
        const InstLayout& L = r.L;
        const size_t xy = (size_t)r.x * r.y;
        if (L.elements() == 0) return;
        BNMF_HIP(hipMemcpyAsync(base + L.off_X, r.X.data(), xy * es_, hipMemcpyHostToDevice, stream_));
        BNMF_HIP(hipMemcpyAsync(base + L.off_alpha, r.alpha.data(), r.x * 8, hipMemcpyHostToDevice, stream_));
        BNMF_HIP(hipMemcpyAsync(base + L.off_beta, r.beta.data(), r.z * 8, hipMemcpyHostToDevice, stream_));
        if (r.has_S0) {
            BNMF_HIP(hipMemcpyAsync(base + L.off_S, r.S0.data(), L.elements() * es_, hipMemcpyHostToDevice,
                                    stream_));
        } else {
            BNMF_HIP(launch_init_s(dtype_, base + L.off_S, base + L.off_X, xy, r.z, r.seed, stream_));
        }
        // the host copies are dropped once the stream has consumed them
        BNMF_HIP(hipStreamSynchronize(stream_));
        r.X.clear();
        r.X.shrink_to_fit();
        r.S0.clear();
        r.S0.shrink_to_fit();
    }


    // Builds descriptors and the work lists of `mode` passes for `set` in the
    // meta buffer region starting at `moff`; returns the end offset.
    size_t build(const std::vector<Request*>& set, WorkSet& ws, std::vector<char>& host, size_t moff,
                 int mode) {
        auto take = [&](size_t n) {
            const size_t o = moff;
            moff = align_up(moff + n);
            if (host.size() < moff) host.resize(moff);
            return o;
        };
        std::vector<BldInst> desc;
        std::vector<PrepItem> prep;
        std::map<PassKey, std::vector<uint32_t>> members;
        std::map<PassKey, uint32_t> zmax;
        std::map<PassKey, size_t> work;
        for (size_t n = 0; n < set.size(); ++n) {
            const Request& r = *set[n];
            const InstLayout& L = r.L;
            BldInst d = make_desc(L, dev_buf_.as<char>(r.off), eps_, r.a, r.b, false, r.nonneg);
            desc.push_back(d);
            if (L.elements() == 0) continue;
            const PassKey key = pass_key(dtype_, mode, L, r.nonneg);
            zmax[key] = std::max(zmax[key], L.z);
            members[key].push_back((uint32_t)n);
            work[key] += L.elements();
            for (uint32_t c = 0; c < L.n_cj; ++c) prep.push_back(PrepItem{(uint32_t)n, 0, c, 0});
            for (uint32_t c = 0; c < L.nA; ++c) prep.push_back(PrepItem{(uint32_t)n, 1, c, 0});
        }
        ws.n_inst = (uint32_t)desc.size();
        ws.off_desc = take(desc.size() * sizeof(BldInst));
        std::memcpy(host.data() + ws.off_desc, desc.data(), desc.size() * sizeof(BldInst));
        ws.n_prep = (uint32_t)prep.size();
        ws.off_prep = take(prep.size() * sizeof(PrepItem));
        std::memcpy(host.data() + ws.off_prep, prep.data(), prep.size() * sizeof(PrepItem));
        ws.off_prep_count = take(4);
        std::memcpy(host.data() + ws.off_prep_count, &ws.n_prep, 4);
        ws.launches.clear();
        std::vector<size_t> load(4, 0);
        std::vector<std::pair<size_t, PassKey>> order;
        for (auto& g : members) order.push_back({work[g.first], g.first});
        std::stable_sort(order.begin(), order.end(), [](auto& a, auto& b) { return a.first > b.first; });
        // update sets: pass items number instances within their stream
        const bool per_lane = mode == MODE_UPDATE;
        std::vector<uint32_t> lane_members[4];
        for (auto& o : order) {
            Launch l;
            l.key = o.second;
            l.zmax = zmax[o.second];
            l.work = o.first;
            l.lane = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[l.lane] += l.work;
            std::vector<PassItem> items;
            for (uint32_t n : members[o.second]) {
                uint32_t idx = n;
                if (per_lane) {
                    idx = (uint32_t)lane_members[l.lane].size();
                    lane_members[l.lane].push_back(n);
                }
                const InstLayout& L = set[n]->L;
                for (uint32_t ti = 0; ti < L.n_ti; ++ti)
                    for (uint32_t tj = 0; tj < L.n_tj; ++tj) items.push_back(PassItem{idx, ti, tj, 0});
            }
            l.n = (uint32_t)items.size();
            l.off_items = take(items.size() * sizeof(PassItem));
            std::memcpy(host.data() + l.off_items, items.data(), items.size() * sizeof(PassItem));
            l.off_count = take(4);
            std::memcpy(host.data() + l.off_count, &l.n, 4);
            ws.launches.push_back(l);
        }
        if (per_lane) {
            for (int ln = 0; ln < 4; ++ln) {
                LaneWS& lw = ws.lanes[ln];
                std::vector<BldInst> ld;
                std::vector<PrepItem> lp;
                for (uint32_t loc = 0; loc < (uint32_t)lane_members[ln].size(); ++loc) {
                    const uint32_t n = lane_members[ln][loc];
                    ld.push_back(desc[n]);
                    const InstLayout& L = set[n]->L;
                    for (uint32_t c = 0; c < L.n_cj; ++c) lp.push_back(PrepItem{loc, 0, c, 0});
                    for (uint32_t c = 0; c < L.nA; ++c) lp.push_back(PrepItem{loc, 1, c, 0});
                }
                lw.n_inst = (uint32_t)ld.size();
                lw.n_prep = (uint32_t)lp.size();
                lw.off_desc = take(ld.size() * sizeof(BldInst));
                if (!ld.empty()) std::memcpy(host.data() + lw.off_desc, ld.data(), ld.size() * sizeof(BldInst));
                lw.off_prep = take(lp.size() * sizeof(PrepItem));
                if (!lp.empty()) std::memcpy(host.data() + lw.off_prep, lp.data(), lp.size() * sizeof(PrepItem));
                lw.off_prep_count = take(4);
                std::memcpy(host.data() + lw.off_prep_count, &lw.n_prep, 4);
            }
        }
        return moff;
    }

    void ensure_meta(size_t bytes) {
        if (meta_.bytes >= bytes) return;
        BNMF_HIP(hipStreamSynchronize(stream_));
        meta_ = DeviceBuffer(std::max(bytes, meta_.bytes * 2));
    }

    void passes(const WorkSet& ws, int mode) {
        char* m = meta_.as<char>();
        const BldInst* d = reinterpret_cast<const BldInst*>(m + ws.off_desc);
        const bool fork = ws.launches.size() > 1;
        if (fork) {
            BNMF_HIP(hipEventRecord(fork_[0], stream_));
            for (int l = 1; l < 4; ++l) BNMF_HIP(hipStreamWaitEvent(side_[l], fork_[0], 0));
        }
        for (const Launch& l : ws.launches) {
            hipStream_t st = fork ? side_[l.lane] : stream_;
            BNMF_HIP(launch_pass_key(l.key, dtype_, psi_, mode, dim3(l.n), st, d,
                                     reinterpret_cast<const PassItem*>(m + l.off_items),
                                     reinterpret_cast<const uint32_t*>(m + l.off_count), l.zmax));
        }
        if (fork) {
            for (int l = 1; l < 4; ++l) {
                BNMF_HIP(hipEventRecord(join_[l], side_[l]));
                BNMF_HIP(hipStreamWaitEvent(stream_, join_[l], 0));
            }
        }
    }

    // iters iterations of an update work set: each stream preps and
    // updates its own instances, iteration after iteration, with one fork
    // and one join (instances on different streams never interact)
    void iterate(const WorkSet& ws, size_t iters) {
        char* m = meta_.as<char>();
        int used = 0;
        for (int l = 0; l < 4; ++l) used += ws.lanes[l].n_inst > 0;
        const bool fork = used > 1;
        if (fork) {
            BNMF_HIP(hipEventRecord(fork_[0], stream_));
            for (int l = 1; l < 4; ++l) BNMF_HIP(hipStreamWaitEvent(side_[l], fork_[0], 0));
        }
        for (size_t t = 0; t < iters; ++t) {
            for (int l = 0; l < 4; ++l) {
                const LaneWS& lw = ws.lanes[l];
                if (lw.n_inst == 0) continue;
                hipStream_t st = fork ? side_[l] : stream_;
                const BldInst* d = reinterpret_cast<const BldInst*>(m + lw.off_desc);
                BNMF_HIP(launch_prep(dtype_, psi_, false, dim3(lw.n_prep), st, d,
                                     reinterpret_cast<const PrepItem*>(m + lw.off_prep),
                                     reinterpret_cast<const uint32_t*>(m + lw.off_prep_count), lw.n_inst));
                for (const Launch& ln : ws.launches) {
                    if (ln.lane != l) continue;
                    BNMF_HIP(launch_pass_key(ln.key, dtype_, psi_, MODE_UPDATE, dim3(ln.n), st, d,
                                             reinterpret_cast<const PassItem*>(m + ln.off_items),
                                             reinterpret_cast<const uint32_t*>(m + ln.off_count), ln.zmax));
                }
            }
        }
        if (fork) {
            for (int l = 1; l < 4; ++l) {
                BNMF_HIP(hipEventRecord(join_[l], side_[l]));
                BNMF_HIP(hipStreamWaitEvent(stream_, join_[l], 0));
            }
        }
    }

    void step(const std::vector<std::shared_ptr<Request>>& act) {
        const double t0 = now_s();
        std::vector<Request*> all, fresh;
        size_t k_iters = max_chunk_;
        for (auto& r : act) {
            all.push_back(r.get());
            if (r->fresh) fresh.push_back(r.get());
            k_iters = std::min(k_iters, r->max_iter - r->done_iter);
        }
        std::vector<char> host;
        WorkSet ws_all, ws_fresh;
        size_t end = build(all, ws_all, host, 0, MODE_UPDATE);
        if (!fresh.empty()) end = build(fresh, ws_fresh, host, end, MODE_SUMS);
        ensure_meta(end);
        BNMF_HIP(hipMemcpyAsync(meta_.ptr, host.data(), end, hipMemcpyHostToDevice, stream_));
        char* m = meta_.as<char>();
        if (!fresh.empty()) passes(ws_fresh, MODE_SUMS);
        iterate(ws_all, k_iters);
        // requests that finish in this step, scored if asked
        std::vector<Request*> fin, score;
        for (Request* r : all) {
            r->fresh = false;
            r->done_iter += k_iters;
            if (r->done_iter >= r->max_iter) {
                fin.push_back(r);
                if (r->want_lm && r->L.elements()) score.push_back(r);
            }
        }
        std::vector<double> lms(score.size(), NAN);
        if (!score.empty()) {
            WorkSet ws_sc;
            std::vector<char> h2;
            const size_t e2 = build(score, ws_sc, h2, 0, MODE_SCORE);
            // a second region after the step's lists, so they stay valid
            const size_t base = align_up(end);
            ensure_meta(base + e2);
            m = meta_.as<char>();
            for (auto& l : ws_sc.launches) {
                l.off_items += base;
                l.off_count += base;
            }
            ws_sc.off_desc += base;
            ws_sc.off_prep += base;
            ws_sc.off_prep_count += base;
            BNMF_HIP(hipMemcpyAsync(m + base, h2.data(), e2, hipMemcpyHostToDevice, stream_));
            passes(ws_sc, MODE_SCORE);
            BNMF_HIP(launch_score(dim3(ws_sc.n_prep), stream_, reinterpret_cast<const BldInst*>(m + ws_sc.off_desc),
                                  reinterpret_cast<const PrepItem*>(m + ws_sc.off_prep),
                                  reinterpret_cast<const uint32_t*>(m + ws_sc.off_prep_count), ws_sc.n_inst));
            for (size_t n = 0; n < score.size(); ++n)
                BNMF_HIP(hipMemcpyAsync(&lms[n], dev_buf_.as<char>(score[n]->off) + score[n]->L.off_result, 8,
                                        hipMemcpyDeviceToHost, stream_));
        }
        BNMF_HIP(hipStreamSynchronize(stream_));
        const double t1 = now_s();
        std::lock_guard<std::mutex> lk(mu_);
        busy_s_ += t1 - t0;
        ++steps_;
        for (size_t n = 0; n < score.size(); ++n) score[n]->log_marginal = lms[n];
        if (!fin.empty()) {
            for (Request* r : fin) {
                r->state = RS_DONE;
                r->t_done = t1;
                ++completed_;
            }
            active_.erase(std::remove_if(active_.begin(), active_.end(),
                                         [](const std::shared_ptr<Request>& r) { return r->state == RS_DONE; }),
                          active_.end());
            done_cv_.notify_all();
        }
    }

    void fail(const std::string& msg) {
        std::lock_guard<std::mutex> lk(mu_);
        failed_ = true;
        fail_msg_ = msg;
        for (auto& kv : all_) {
            auto& r = kv.second;
            if (r->state < RS_DONE) {
                r->state = RS_FAILED;
                r->status = BNMF_ERR_DEVICE;
                r->err = msg;
                r->t_done = now_s();
            }
        }
        queue_.clear();
        active_.clear();
        done_cv_.notify_all();
    }

    int dtype_, psi_;
    double eps_;
    size_t es_;
    size_t max_active_;
    size_t max_chunk_ = 8;  // iterations per scheduling step
    int dev_ = 0;
    Arena arena_;
    DeviceBuffer dev_buf_, meta_;
    hipStream_t stream_ = nullptr;
    std::vector<hipStream_t> side_;
    std::vector<hipEvent_t> fork_, join_;
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<std::shared_ptr<Request>> queue_;
    std::vector<std::shared_ptr<Request>> active_;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> all_;
    uint64_t next_ticket_ = 0;
    size_t submitted_ = 0, completed_ = 0, rejected_ = 0, steps_ = 0;
    size_t peak_bytes_ = 0, peak_active_ = 0;
    double busy_s_ = 0;
    bool stop_ = false, failed_ = false;
    std::string fail_msg_;
};

}  // namespace bnmf

// ------------------------------------------------------------------ C ABI
struct bnmf_server {
    std::unique_ptr<bnmf::Server> impl;
};

namespace {
template <typename F>
bnmf_status guarded_s(F&& f) {
    try {
        bnmf::clear_last_error();
        f();
        return BNMF_OK;
    } catch (const bnmf::Error& e) {
        bnmf::set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        bnmf::set_last_error("host allocation failed");
        return BNMF_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        bnmf::set_last_error(e.what());
        return BNMF_ERR_DEVICE;
    }
}
}  // namespace

extern "C" {

bnmf_status bnmf_server_create(bnmf_server** out, int dtype, size_t hbm_budget_bytes,
                               size_t max_active, int use_psi_appr, double eps) {
    return guarded_s([&] {
        BNMF_REQUIRE(out != nullptr, "out is null");
        *out = nullptr;
        BNMF_REQUIRE(dtype == BNMF_F32 || dtype == BNMF_F64, "dtype must be BNMF_F32 or BNMF_F64");
        const int psi = dtype == BNMF_F32 ? bnmf::PSIM_F32
                                          : (use_psi_appr ? bnmf::PSIM_F64_APPR : bnmf::PSIM_F64_EXACT);
        auto s = std::make_unique<bnmf_server>();
        s->impl = std::make_unique<bnmf::Server>(dtype, hbm_budget_bytes, max_active, psi, eps);
        *out = s.release();
    });
}

bnmf_status bnmf_server_submit(bnmf_server* s, const bnmf_bld_desc* req, size_t max_iter,
                               int want_log_marginal, uint64_t* ticket) {
    return guarded_s([&] {
        BNMF_REQUIRE(s && req && ticket, "null argument");
        BNMF_REQUIRE(req->x * req->y == 0 || req->X != nullptr, "X is null");
        BNMF_REQUIRE((req->x == 0 || req->alpha != nullptr) && (req->z == 0 || req->beta != nullptr),
                     "alpha or beta is null");
        BNMF_REQUIRE(req->x < (1u << 31) && req->y < (1u << 31) && req->z < (1u << 31),
                     "x, y and z must be below 2^31");
        *ticket = s->impl->submit(*req, max_iter, want_log_marginal);
    });
}

bnmf_status bnmf_server_poll(bnmf_server* s, uint64_t ticket, int* done) {
    return guarded_s([&] {
        BNMF_REQUIRE(s && done, "null argument");
        bnmf_status st = BNMF_OK;
        *done = s->impl->poll(ticket, &st) ? 1 : 0;
    });
}

bnmf_status bnmf_server_wait(bnmf_server* s, uint64_t ticket, double timeout_s) {
    bnmf_status inner = BNMF_OK;
    const bnmf_status st = guarded_s([&] {
        BNMF_REQUIRE(s != nullptr, "null argument");
        inner = s->impl->wait(ticket, timeout_s);
    });
    return st != BNMF_OK ? st : inner;
}

bnmf_status bnmf_server_collect(bnmf_server* s, uint64_t ticket, void* S_out,
                                double* log_marginal_out, bnmf_request_stats* stats) {
    return guarded_s([&] {
        BNMF_REQUIRE(s != nullptr, "null argument");
        s->impl->collect(ticket, S_out, log_marginal_out, stats);
    });
}

bnmf_status bnmf_server_get_stats(bnmf_server* s, bnmf_server_stats* out) {
    return guarded_s([&] {
        BNMF_REQUIRE(s && out, "null argument");
        *out = s->impl->stats();
    });
}

size_t bnmf_server_request_bytes(int dtype, size_t x, size_t y, size_t z) {
    if (x >= (1u << 31) || y >= (1u << 31) || z >= (1u << 31)) return 0;
    return bnmf::Server::request_bytes(dtype, (uint32_t)x, (uint32_t)y, (uint32_t)z);
}

void bnmf_server_destroy(bnmf_server* s) { delete s; }

}  // extern "C"
