// serve.cpp -- continuous batching of threading requests (see serve.h).
#include "serve.h"

#include <algorithm>
#include <chrono>
#include <cstdio>

namespace aw {

static double wall_seconds()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

Server::Server(const ServeOptions &opt) : opt_(opt), t0_(wall_seconds())
{
    if (opt_.max_batch < 1) opt_.max_batch = 1;
    opt_.drivers = std::max(1, std::min(opt_.drivers, 4));
    for (int i = 0; i < opt_.drivers; i++) drivers_.emplace_back(&Server::worker, this, i);
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (std::thread &t : drivers_) t.join();
    for (Request *r : queue_) delete r;
}

double Server::now() const { return wall_seconds() - t0_; }

// Mirrors the device arrays Batch::run allocates for one instance: the
// forward table, the lane-run planes of every block, the variant rows (the
// variant columns are counted here on the host), the per-block transition
// tables and the per-column arrays.
size_t Server::footprint(const HostInstance &hi, Precision prec)
{
    const FlatTrees &ft = hi.ft;
    const int T = hi.model.T, T1 = T - 1, B = ft.nblocks, N = ft.nnodes, L = ft.seqlen;
    std::vector<int> Sv;
    instance_states(ft, T, &Sv);
    int Smax = 1;
    long long Stot = 0;
    for (int S : Sv) {
        Smax = std::max(Smax, S);
        Stot += S;
    }
    // states per lane as the kernels round it (1, 2, 4, 6, 8, 12, 16)
    size_t R = 16;
    for (int r : {1, 2, 4, 6, 8, 12, 16})
        if (64 * r >= Smax) {
            R = r;
            break;
        }
    const size_t RW = 64 * R;
    long long nvar = 0;
    if (!hi.rows.empty()) {
        for (int c = 0; c < L; c++) {
            const char x = hi.rows[0][c];
            for (size_t r = 1; r < hi.rows.size(); r++)
                if (hi.rows[r][c] != x) {
                    nvar++;
                    break;
                }
        }
    }
    const size_t rsz = prec == Precision::FP32 ? 4 : 8;
    const size_t TC = (size_t)tc_base(T1, T1);
    size_t b = (size_t)L * RW * rsz;                          // forward table
    b += (size_t)B * RW * (NCOEF * rsz + 4 + 4);              // planes, info, targets
    b += (size_t)std::max<long long>(nvar, 1) * RW * 8;        // variant rows
    b += (size_t)B * ((size_t)T1 * T1 + TC) * 8;               // transition tables
    b += (size_t)Stot * (4 + 4 + 48);                          // state tables, emission terms
    b += (size_t)B * N * (5 * 4 + 16 + 4) + (size_t)B * 12;    // trees, pruning
    b += (size_t)L * (3 * 4 + 8 + 4 + 8 + hi.rows.size());     // columns, uniforms, sequences
    return b + (size_t)16 * 4096;
}

int Server::submit(const HostInstance &hi, const double *u)
{
    // a request the kernels cannot run fails on its own, at once, instead of
    // failing the batch it would join
    std::string why;
    if (instance_supported(hi, &why)) {
        fprintf(stderr, "argweaver_amd: request refused: %s\n", why.c_str());
        std::lock_guard<std::mutex> lk(mu_);
        ServeResult res;
        const int id = next_id_++;
        res.id = id;
        res.status = kRefused;
        res.t_submit = res.t_start = res.t_done = now();
        done_[id] = std::move(res);
        cv_done_.notify_all();
        return id;
    }
    Request *r = new Request;
    r->hi = hi;
    const int L = hi.ft.seqlen;
    r->rows.resize(hi.rows.size());
    for (size_t i = 0; i < hi.rows.size(); i++) r->rows[i].assign(hi.rows[i], hi.rows[i] + L);
    for (size_t i = 0; i < hi.rows.size(); i++) r->hi.rows[i] = r->rows[i].data();
    r->u.assign(u, u + L);
    r->bytes = footprint(r->hi, opt_.run.prec);
    r->t_submit = now();
    std::lock_guard<std::mutex> lk(mu_);
    r->id = next_id_++;
    queue_.push_back(r);
    cv_work_.notify_one();
    return r->id;
}

int Server::wait(int id, ServeResult *out, double timeout_s)
{
    std::unique_lock<std::mutex> lk(mu_);
    if (id < 0 || id >= next_id_) return 1;
    auto ready = [&] { return done_.count(id) > 0; };
    if (timeout_s > 0) {
        if (!cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return 2;
    } else {
        cv_done_.wait(lk, ready);
    }
    if (out) *out = std::move(done_[id]);
    done_.erase(id);
    return 0;
}

void Server::worker(int index)
{
    // each driver's batch takes its own device lock, so the drivers' batches
    // run side by side on the device
    Batch batch;
    batch.set_concurrent(true);
    const size_t budget = opt_.hbm_budget / opt_.drivers;
    std::vector<Request *> mine;
    while (true) {
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_work_.wait(lk, [&] { return stop_ || !queue_.empty(); });
            if (queue_.empty()) return;   // stopping, nothing left
            // admission in arrival order while the batch fits the budget
            size_t taken = 0;
            while (!queue_.empty() && (int)mine.size() < opt_.max_batch) {
                const size_t b = queue_.front()->bytes;
                if (!mine.empty() && taken + b > budget) break;
                taken += b;
                mine.push_back(queue_.front());
                queue_.pop_front();
            }
        }
        const double t_start = now();
        batch.clear();
        for (Request *r : mine) batch.add(r->hi, r->u.data());
        int rc = batch.upload();
        if (!rc) rc = batch.run(opt_.run);
        std::vector<ServeResult> res(mine.size());
        for (size_t i = 0; i < mine.size(); i++) {
            ServeResult &o = res[i];
            const Request *r = mine[i];
            o.id = r->id;
            o.status = rc;
            o.t_submit = r->t_submit;
            o.t_start = t_start;
            o.batch_size = (int)mine.size();
            o.driver = index;
            o.W = batch.W((int)i);
            if (!rc) {
                o.lnl = batch.lnl((int)i);
                o.path.resize(r->hi.ft.seqlen);
                if (batch.copy_path((int)i, o.path.data())) o.status = 1;
            }
        }
        const double t_done = now();
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (auto &o : res) {
                o.t_done = t_done;
                done_[o.id] = std::move(o);
            }
        }
        for (Request *r : mine) delete r;
        mine.clear();
        cv_done_.notify_all();
    }
}

}  // namespace aw
