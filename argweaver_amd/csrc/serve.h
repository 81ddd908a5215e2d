// serve.h -- continuous batching of threading requests on one GPU.
//
// Requests (one threading instance each: the ARG with one lineage removed,
// the sequences, the uniforms of the traceback) are submitted from any host
// thread.  Driver threads own the device work, each with a batch of its own:
// a driver that is free takes every queued request, in arrival order, whose
// estimated HBM footprint still fits its share of the budget (a request
// larger than the share runs alone) and runs them as one batch.  With two
// drivers, requests that arrive while a batch is on the GPU are admitted by
// the other driver at once and their work overlaps on the device with the
// batch in flight.  All HIP calls of a server come from its driver threads.
// Each request's result does not depend on the batch it ran in.
#pragma once

#include <condition_variable>
#include <cstddef>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine.h"

namespace aw {

struct ServeOptions {
    size_t hbm_budget = size_t(64) << 30;  // bytes of device memory, shared by the drivers
    int max_batch = 64;                    // requests per batch
    int drivers = 2;                       // batches in flight at once (1 .. 4)
    BatchOptions run;                      // precision and traceback tiles
};

// status of a request the kernels cannot run (over the node, time-point or
// states-per-block limits), refused at submission
constexpr int kRefused = 3;

struct ServeResult {
    int id = -1;
    int status = 0;          // 0 ok, kRefused, else the batch failed
    double lnl = 0.0;
    std::vector<int> path;
    long long W = 0;         // column-states of the request
    double t_submit = 0, t_start = 0, t_done = 0;   // seconds on the server clock
    int batch_size = 0;
    int driver = 0;          // the driver thread that ran it
};

class Server {
public:
    explicit Server(const ServeOptions &opt);
    ~Server();
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;

    // Takes a copy of the sequence rows and the uniforms.  Returns the id.
    int submit(const HostInstance &hi, const double *u);
    // Blocks until request id is done (at most timeout_s seconds when
    // timeout_s > 0); the result is removed from the server.  Returns 0, 1
    // for an unknown id, 2 on time-out (the request stays pending).
    int wait(int id, ServeResult *out, double timeout_s = 0);
    // Seconds since the server started (the clock of the result times).
    double now() const;
    // Device bytes a request needs while it runs (the admission estimate).
    static size_t footprint(const HostInstance &hi, Precision prec);

private:
    struct Request {
        int id;
        HostInstance hi;
        std::vector<std::string> rows;   // owned copies of the sequence rows
        std::vector<double> u;
        size_t bytes;
        double t_submit;
    };
    void worker(int index);

    ServeOptions opt_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<Request *> queue_;
    std::map<int, ServeResult> done_;
    int next_id_ = 0;
    bool stop_ = false;
    std::vector<std::thread> drivers_;
    double t0_;
};

}  // namespace aw
