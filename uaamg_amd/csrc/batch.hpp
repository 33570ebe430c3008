// batch.hpp -- many independent pressure solves in one set of launches.
//
// Each member grid keeps its own device hierarchy (an Engine built and set up
// as for a single solve); the batch adds one descriptor per grid, one Ctrl
// slot per grid and, per phase of the PCG iteration, a work-list of
// (grid, tile) items.  One PCG iteration of the whole batch is then a fixed
// sequence of about 15 launches whatever the number of members, which the
// batch captures once into a hipGraph and replays.  Members that converge skip
// their items; the batch polls the done flags between replays.
//
// Launch shapes (the march kernels' plane chunk per level) come from a small
// autotuner: on the first solve after the membership changes, each candidate
// work-list density is timed on a few V-cycles and the fastest is kept.
#pragma once
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <vector>

#include "engine.hpp"

namespace uaamg {

// what the C-ABI sees of a batch of one dimension and level precision
class BatchBase {
public:
    virtual ~BatchBase() = default;
    virtual int add(const int n[3], const double h[3], const int wall[6], const uint8_t* solid_dev) = 0;
    virtual void set_mask(int i, const uint8_t* solid_dev) = 0;
    virtual void clear() = 0;
    virtual int size() const = 0;
    virtual size_t bytes() const = 0;
    virtual void begin(const double* const* rhs, double* const* pout, const int* warm, int max_iter, double tol) = 0;
    virtual bool done() const = 0;
    virtual void chunk() = 0;
    virtual void poll() = 0;
    virtual void finish() = 0;
    virtual void stats(int g, uaamg_stats* st, double ms_total) const = 0;
    virtual int tuned_target() const = 0;
    virtual int tuned_kc_min() const = 0;
    virtual int tuned_rows_kc() const = 0;
    virtual void prepare() = 0;
    virtual void start_member(int g, const double* rhs, double* pout, int warm, int max_iter, double tol) = 0;
    virtual void finish_member(int g, double* pout, int warm) = 0;
    virtual void run_chunk() = 0;
    virtual const Ctrl& member_ctrl(int g) const = 0;
    virtual void remove(int g) = 0;
    virtual size_t member_bytes(int g) const = 0;
};

template <int D, typename S>
class BatchEngine final : public BatchBase {
    using E = Engine<D, float, S>;
    using C4 = typename Vec4<S>::type;
    static constexpr int RX0 = E::RX0, RY0 = E::RY0, PPT0 = E::PPT0;
    static constexpr int RXC = E::RXC, RYC = E::RYC, PPTC = E::PPTC;
    static constexpr int NLMAX = E::TAIL_NLMAX;

public:
    BatchEngine(const uaamg_opts& o, hipStream_t s)
        : o_(o), s_(s), kc_min_(std::max(2, o.march_kc_min)), rows_kc_(std::max(0, o.rows_kc)) {}
    ~BatchEngine() {
        destroy_graph();
        free_all();
    }
    BatchEngine(const BatchEngine&) = delete;
    BatchEngine& operator=(const BatchEngine&) = delete;

    int add(const int n[3], const double h[3], const int wall[6], const uint8_t* solid_dev) override {
        // members run the batched kernels only: no fused p-update buffers or
        // x-update stream of the single-grid form
        uaamg_opts mo = o_;
        mo.forms &= ~UAAMG_FORM_PMV;
        auto e = std::make_unique<E>(mo, n, h, wall, s_);
        e->set_grid(solid_dev);
        e->set_idle();
        if (!e->batch_ready())
            throw HipError{hipErrorInvalidValue, "grid cannot join a batch (needs V(1,1), >= 2 levels and a tail level)"};
        eng_.push_back(std::move(e));
        dirty_ = true;
        return (int)eng_.size() - 1;
    }
    void set_mask(int i, const uint8_t* solid_dev) override {
        eng_[i]->set_grid(solid_dev);
        dirty_ = true;
    }
    void clear() override {
        destroy_graph();
        eng_.clear();
        dirty_ = true;
    }
    int size() const override { return (int)eng_.size(); }
    size_t bytes() const override {
        size_t b = 0;
        for (auto& e : eng_) b += e->bytes();
        return b + bytes_;
    }
    void set_stream(hipStream_t s) {
        s_ = s;
        for (auto& e : eng_) e->set_stream(s);
        destroy_graph();
    }

    // ---- solve, in phases so that a caller can interleave several batches ----
    void begin(const double* const* rhs, double* const* pout, const int* warm, int max_iter, double tol) override {
        if (eng_.empty()) return;
        if (dirty_) build();
        const int G = size();
        any_warm_ = false;
        for (int g = 0; g < G; g++) {
            hd_[g].rhs = rhs[g];
            hd_[g].pout = pout[g];
            hd_[g].warm = warm ? warm[g] : 0;
            any_warm_ |= hd_[g].warm != 0;
        }
        HIPCHK(hipMemcpyAsync(dd_, hd_.data(), sizeof(BGrid<S>) * G, hipMemcpyHostToDevice, s_));
        max_iter_ = max_iter;
        queued_ = 0;
        done_ = false;
        kb_ctrl_init<S><<<(G + 63) / 64, 64, 0, s_>>>(dd_, G, max_iter, tol, o_.tol_mode, o_.flexible, o_.project_mean);
        kb_cast_warm<D, S><<<nflat_, 256, 0, s_>>>(dd_, flat_);
        kb_init_sum<S><<<nflat_, 256, 0, s_>>>(dd_, flat_, pflat_);
        kb_fin_init_sum<S><<<G, FIN_THREADS, 0, s_>>>(dd_, pflat_, nflat_);
        kb_init_r<S><<<nflat_, 256, 0, s_>>>(dd_, flat_, pflat_);
        kb_fin_init_r<S><<<G, FIN_THREADS, 0, s_>>>(dd_, pflat_, nflat_);
        kb_init_sub<S><<<nflat_, 256, 0, s_>>>(dd_, flat_, pflat_);
        kb_fin_init_sub<S><<<G, FIN_THREADS, 0, s_>>>(dd_, pflat_, nflat_);
        if (any_warm_) {
            kb_init_warm<D, S><<<nflat_, 256, 0, s_>>>(dd_, flat_, pflat_);
            kb_fin_init_warm<S><<<G, FIN_THREADS, 0, s_>>>(dd_, pflat_, nflat_);
        }
        vcycle();
        kb_pupdate<S><<<nflat_, 256, 0, s_>>>(dd_, flat_);
        HIPCHK(hipGetLastError());
    }
    bool done() const override { return done_ || eng_.empty(); }
    // enqueue the next chunk of iterations and the copy of the scalars
    void chunk() override {
        if (done()) return;
        const int every = std::max(1, o_.check_every);
        const int n = std::min(every, std::max(1, max_iter_ - queued_));
        if (o_.use_graph && n == every) {
            if (!gexec_) {
                HIPCHK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
                for (int c = 0; c < n; c++) iteration();
                HIPCHK(hipStreamEndCapture(s_, &graph_));
                HIPCHK(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
            }
            HIPCHK(hipGraphLaunch(gexec_, s_));
        } else {
            for (int c = 0; c < n; c++) iteration();
        }
        queued_ += n;
        copy_ctrl();
    }
    void copy_ctrl() {
        kb_gather_ctrl<S><<<(size() + 63) / 64, 64, 0, s_>>>(dd_, size(), ctrl_);
        HIPCHK(hipMemcpyAsync(hctrl_, ctrl_, sizeof(Ctrl) * size(), hipMemcpyDeviceToHost, s_));
    }
    // after the stream has been synchronised
    void poll() override {
        if (done()) return;
        bool all = true;
        for (int g = 0; g < size(); g++) all = all && hctrl_[g].done;
        done_ = all || queued_ >= max_iter_;
    }
    void finish() override {
        if (eng_.empty()) return;
        kb_finalize<S><<<nflat_, 256, 0, s_>>>(dd_, flat_);
        copy_ctrl();
        HIPCHK(hipGetLastError());
    }
    // after finish() and a synchronisation
    void stats(int g, uaamg_stats* st, double ms_total) const override {
        const Ctrl& c = hctrl_[g];
        std::memset(st, 0, sizeof(*st));
        st->iters = c.iter;
        st->status = c.status == ST_CONVERGED ? UAAMG_OK
                   : c.status == ST_BREAKDOWN ? UAAMG_BREAKDOWN
                   : c.status == ST_NONFINITE ? UAAMG_NONFINITE : UAAMG_NOT_CONVERGED;
        st->nlevels = eng_[g]->nlevels();
        st->rnorm = std::sqrt(std::max(0.0, c.rsnew));
        st->r0norm = std::sqrt(std::max(0.0, c.r0sq));
        st->ms_total = ms_total;
        st->alg_bytes = eng_[g]->alg_bytes_of(c.iter);
    }
    int tuned_target() const override { return target_; }
    int tuned_kc_min() const override { return kc_min_; }
    int tuned_rows_kc() const override { return rows_kc_; }

    // ---- continuous batching: members start and finish on their own ----
    // (used by the server: members join between chunks, converged members
    // leave; idle members skip every batched launch)
    void prepare() override {
        if (dirty_) build();
    }
    void start_member(int g, const double* rhs, double* pout, int warm, int max_iter, double tol) override {
        eng_[g]->start(rhs, pout, warm, max_iter, tol);
    }
    void finish_member(int g, double* pout, int warm) override { eng_[g]->finalize_into(pout, warm); }
    void run_chunk() override {
        if (eng_.empty()) return;
        if (dirty_) build();
        const int every = std::max(1, o_.check_every);
        if (o_.use_graph) {
            if (!gexec_) {
                HIPCHK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
                for (int c = 0; c < every; c++) iteration();
                HIPCHK(hipStreamEndCapture(s_, &graph_));
                HIPCHK(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
            }
            HIPCHK(hipGraphLaunch(gexec_, s_));
        } else {
            for (int c = 0; c < every; c++) iteration();
        }
        copy_ctrl();
    }
    const Ctrl& member_ctrl(int g) const override { return hctrl_[g]; }
    void remove(int g) override {
        HIPCHK(hipStreamSynchronize(s_));
        destroy_graph();
        eng_.erase(eng_.begin() + g);
        dirty_ = true;
    }
    size_t member_bytes(int g) const override { return eng_[g]->bytes(); }

private:
    // ---- one PCG iteration of every member ----
    void iteration() {
        const int G = size();
        kb_matvec<D, S><<<nflat_, 256, 0, s_>>>(dd_, flat_, pflat_);
        kb_fin_matvec<S><<<G, FIN_THREADS, 0, s_>>>(dd_, pflat_, nflat_);
        if (o_.flexible) {
            kb_update<S, true><<<nflat_, 256, 0, s_>>>(dd_, flat_, pflat_);
            kb_fin_update<S, true><<<G, FIN_THREADS, 0, s_>>>(dd_, pflat_, nflat_);
        } else {
            kb_update<S, false><<<nflat_, 256, 0, s_>>>(dd_, flat_, pflat_);
            kb_fin_update<S, false><<<G, FIN_THREADS, 0, s_>>>(dd_, pflat_, nflat_);
        }
        vcycle();
        kb_pupdate<S><<<nflat_, 256, 0, s_>>>(dd_, flat_);
    }
    void vcycle() {
        const int G = size();
        if constexpr (D == 3) {
            if (nrows_[0]) kb_rows_down0<2, S><<<nrows_[0], 640, 0, s_>>>(dd_, rows_[0]);
            if (nrows_[1]) kb_rows_down0<4, S><<<nrows_[1], 640, 0, s_>>>(dd_, rows_[1]);
        }
        if (nm0_) kb_march_down0<D, RX0, RY0, PPT0, S><<<nm0_, 256, 0, s_>>>(dd_, m0_);
        for (int l = 1; l < (int)nc_.size(); l++)
            if (nc_[l]) kb_march_down<D, RXC, RYC, PPTC, S><<<nc_[l], 256, 0, s_>>>(dd_, c_[l], l);
        kb_tail<D, S, NLMAX><<<G, TAIL_THREADS, 0, s_>>>(dd_);
        for (int l = (int)nc_.size() - 1; l >= 1; l--)
            if (nc_[l]) kb_march_up<D, RXC, RYC, PPTC, S><<<nc_[l], 256, 0, s_>>>(dd_, c_[l], l);
        if constexpr (D == 3) {
            if (nrows_[0]) kb_rows_up0<2, S><<<nrows_[0], 640, 0, s_>>>(dd_, rows_[0], prows_[0]);
            if (nrows_[1]) kb_rows_up0<4, S><<<nrows_[1], 640, 0, s_>>>(dd_, rows_[1], prows_[1]);
        }
        if (nm0_) kb_march_up0<D, RX0, RY0, PPT0, S><<<nm0_, 256, 0, s_>>>(dd_, m0_, pm0_);
        kb_fin_zsums<S><<<G, FIN_THREADS, 0, s_>>>(dd_);
    }

    // ---- descriptors and work-lists ----
    void build() {
        destroy_graph();
        free_all();
        const int G = size();
        if (G == 0) { dirty_ = false; return; }
        ctrl_ = dev<Ctrl>(G);
        HIPCHK(hipHostMalloc((void**)&hctrl_, sizeof(Ctrl) * G, hipHostMallocDefault));
        std::memset(hctrl_, 0, sizeof(Ctrl) * G);
        dd_ = dev<BGrid<S>>(G);
        keep_ = lists_.size();   // ctrl and descriptors stay; work-lists are rebuilt by the autotuner
        hd_.assign(G, BGrid<S>{});
        for (int g = 0; g < G; g++) eng_[g]->fill_batch(hd_[g], eng_[g]->ctrl_ptr());
        make_lists(target_);
        dirty_ = false;
        if (!tuned_) autotune();
    }

    void make_lists(long target) {
        const int G = size();
        std::vector<BItem> flat, rows[2], m0;
        std::vector<std::vector<BItem>> cl;
        for (int g = 0; g < G; g++) {
            BGrid<S>& d = hd_[g];
            // flat phases: 4 cells x 256 threads x 4 loop trips per workgroup
            const long nb = std::min<long>(2048, std::max<long>(1, (d.N4 / 4 + 1023) / 1024));
            d.flat_base = (int)flat.size();
            d.flat_cnt = (int)nb;
            for (long b = 0; b < nb; b++) flat.push_back(BItem{g, (int)b, (int)nb, 0});
            // level 0
            if (D == 3 && d.cpl) {
                // row tiles of rows_kc_ planes (0: the engine's own choice)
                d.rgeo = E::make_rowgeo(d.g0.ny, d.g0.nz, E::ROWS_NW * E::ROWS_RW - 4, rows_kc_);
                std::vector<BItem>& L = rows[d.cpl == 2 ? 0 : 1];
                d.up_base = (int)L.size();
                d.up_cnt = d.rgeo.ntiles;
                for (int t = 0; t < d.rgeo.ntiles; t++) L.push_back(BItem{g, t, 0, 0});
            } else {
                typename E::MLaunch m = E::march_launch(d.g0, RX0, RY0, target, kc_min_);
                d.lKC[0] = m.KC;
                d.up_base = (int)m0.size();
                d.up_cnt = (int)(m.grid.x * m.grid.y * m.grid.z);
                for (unsigned z = 0; z < m.grid.z; z++)
                    for (unsigned y = 0; y < m.grid.y; y++)
                        for (unsigned x = 0; x < m.grid.x; x++) m0.push_back(BItem{g, (int)x, (int)y, (int)z});
            }
            // march levels 1 .. lt-1
            for (int l = 1; l < d.lt; l++) {
                if ((int)cl.size() <= l) cl.resize(l + 1);
                typename E::MLaunch m = E::march_launch(d.lg[l], RXC, RYC, target, kc_min_);
                d.lKC[l] = m.KC;
                for (unsigned z = 0; z < m.grid.z; z++)
                    for (unsigned y = 0; y < m.grid.y; y++)
                        for (unsigned x = 0; x < m.grid.x; x++) cl[l].push_back(BItem{g, (int)x, (int)y, (int)z});
            }
        }
        nflat_ = (int)flat.size();
        flat_ = upload(flat);
        pflat_ = dev<double>(3 * (size_t)nflat_ + 64);
        for (int c = 0; c < 2; c++) {
            nrows_[c] = (int)rows[c].size();
            rows_[c] = nrows_[c] ? upload(rows[c]) : nullptr;
            prows_[c] = nrows_[c] ? dev<double>(3 * (size_t)nrows_[c] + 64) : nullptr;
        }
        nm0_ = (int)m0.size();
        m0_ = nm0_ ? upload(m0) : nullptr;
        pm0_ = nm0_ ? dev<double>(3 * (size_t)nm0_ + 64) : nullptr;
        nc_.assign(cl.size(), 0);
        c_.assign(cl.size(), nullptr);
        for (size_t l = 1; l < cl.size(); l++) {
            nc_[l] = (int)cl[l].size();
            c_[l] = nc_[l] ? upload(cl[l]) : nullptr;
        }
        for (int g = 0; g < G; g++) {
            BGrid<S>& d = hd_[g];
            if (D == 3 && d.cpl) {
                const int c = d.cpl == 2 ? 0 : 1;
                d.up_partials = prows_[c];
                d.up_nbt = nrows_[c];
            } else {
                d.up_partials = pm0_;
                d.up_nbt = nm0_;
            }
        }
        HIPCHK(hipMemcpyAsync(dd_, hd_.data(), sizeof(BGrid<S>) * G, hipMemcpyHostToDevice, s_));
        HIPCHK(hipStreamSynchronize(s_));
    }

    // Autotuner over the launch shapes, by coordinate descent: the march
    // work-list density (workgroups aimed at per level and grid: 64, 256,
    // 1024), then the fewest planes per march chunk (2, 4, 8), then the planes
    // per level-0 row tile of 3D members (the engine's choice, 8, 16).  Each
    // candidate is timed on a few V-cycles of the real members (zero right-
    // hand sides: the kernels run all their arithmetic regardless) and the
    // fastest is kept for the batch.  Every candidate computes the same
    // arithmetic; only the tiling changes.
    void autotune() {
        tuned_ = true;
        const int G = size();
        bool any_rows = false;
        for (int g = 0; g < G; g++) any_rows = any_rows || (D == 3 && hd_[g].cpl);
        std::vector<BGrid<S>> save = hd_;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        auto time_it = [&]() {
            free_lists();
            make_lists(target_);
            kb_ctrl_init<S><<<(G + 63) / 64, 64, 0, s_>>>(dd_, G, 1 << 30, 0.0, o_.tol_mode, 0, o_.project_mean);
            vcycle();   // warm-up
            HIPCHK(hipEventRecord(e0, s_));
            for (int k = 0; k < 3; k++) vcycle();
            HIPCHK(hipEventRecord(e1, s_));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            return ms;
        };
        float best = 1e30f;
        long bt = target_;
        for (long t : {64L, 256L, 1024L}) {
            target_ = t;
            const float ms = time_it();
            if (ms < best) { best = ms; bt = t; }
        }
        target_ = bt;
        int bk = kc_min_;
        for (int k : {4, 8}) {
            kc_min_ = k;
            const float ms = time_it();
            if (ms < best) { best = ms; bk = k; }
        }
        kc_min_ = bk;
        if (any_rows) {
            int br = rows_kc_;
            for (int r : {8, 16}) {
                rows_kc_ = r;
                const float ms = time_it();
                if (ms < best) { best = ms; br = r; }
            }
            rows_kc_ = br;
        }
        HIPCHK(hipEventDestroy(e0));
        HIPCHK(hipEventDestroy(e1));
        free_lists();
        hd_ = save;
        make_lists(target_);
        kb_ctrl_init<S><<<(G + 63) / 64, 64, 0, s_>>>(dd_, G, 0, 0.0, o_.tol_mode, 0, o_.project_mean);   // idle again
        HIPCHK(hipStreamSynchronize(s_));
    }

    void destroy_graph() {
        if (gexec_) (void)hipGraphExecDestroy(gexec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        gexec_ = nullptr;
        graph_ = nullptr;
    }
    template <typename P> P* dev(size_t n) {
        void* p = nullptr;
        HIPCHK(hipMalloc(&p, std::max<size_t>(8, n * sizeof(P))));
        HIPCHK(hipMemsetAsync(p, 0, std::max<size_t>(8, n * sizeof(P)), s_));
        bytes_ += n * sizeof(P);
        lists_.push_back(p);
        return reinterpret_cast<P*>(p);
    }
    template <typename P> P* upload(const std::vector<P>& v) {
        P* d = dev<P>(v.size());
        HIPCHK(hipMemcpyAsync(d, v.data(), sizeof(P) * v.size(), hipMemcpyHostToDevice, s_));
        return d;
    }
    // work-lists and their partials (everything allocated after dd_)
    void free_lists() {
        HIPCHK(hipStreamSynchronize(s_));
        while (lists_.size() > keep_) {
            (void)hipFree(lists_.back());
            lists_.pop_back();
        }
        flat_ = nullptr; m0_ = nullptr; pflat_ = nullptr; pm0_ = nullptr;
        for (int c = 0; c < 2; c++) { rows_[c] = nullptr; prows_[c] = nullptr; nrows_[c] = 0; }
        nflat_ = nm0_ = 0;
        nc_.clear(); c_.clear();
    }
    void free_all() {
        if (!lists_.empty() || hctrl_) (void)hipStreamSynchronize(s_);
        for (void* p : lists_) (void)hipFree(p);
        lists_.clear();
        keep_ = 0;
        if (hctrl_) (void)hipHostFree(hctrl_);
        hctrl_ = nullptr; ctrl_ = nullptr; dd_ = nullptr;
        flat_ = nullptr; m0_ = nullptr; pflat_ = nullptr; pm0_ = nullptr;
        for (int c = 0; c < 2; c++) { rows_[c] = nullptr; prows_[c] = nullptr; nrows_[c] = 0; }
        nflat_ = nm0_ = 0;
        nc_.clear(); c_.clear();
        bytes_ = 0;
    }

    uaamg_opts o_;
    hipStream_t s_;
    std::vector<std::unique_ptr<E>> eng_;
    bool dirty_ = true, tuned_ = false, done_ = true, any_warm_ = false;
    long target_ = 256;
    int kc_min_ = 2;      // fewest planes per march chunk
    int rows_kc_ = 0;     // planes per level-0 row tile (0: the engine's choice)
    int max_iter_ = 0, queued_ = 0;
    std::vector<BGrid<S>> hd_;
    BGrid<S>* dd_ = nullptr;
    Ctrl* ctrl_ = nullptr;
    Ctrl* hctrl_ = nullptr;
    BItem *flat_ = nullptr, *m0_ = nullptr, *rows_[2] = {nullptr, nullptr};
    double *pflat_ = nullptr, *pm0_ = nullptr, *prows_[2] = {nullptr, nullptr};
    int nflat_ = 0, nm0_ = 0, nrows_[2] = {0, 0};
    std::vector<int> nc_;
    std::vector<BItem*> c_;
    std::vector<void*> lists_;
    size_t keep_ = 0;
    size_t bytes_ = 0;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t gexec_ = nullptr;
};

}  // namespace uaamg
