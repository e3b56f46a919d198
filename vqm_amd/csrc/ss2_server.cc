// Request server with continuous batching (see ss2_server.h).

#include "ss2_server.h"

#include "ss2_source.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>

namespace vqm::ss2 {

uint64_t NowNs() {
    return static_cast<uint64_t>(std::chrono::duration_cast<std::chrono::nanoseconds>(
                                     std::chrono::steady_clock::now().time_since_epoch())
                                     .count());
}

namespace {

const vqm_ss2_server_config& ValidatedServer(const vqm_ss2_server_config& c) {
    ToFormat(c.plane_format);
    if (c.lanes > 16) Fail(VQM_INVALID_ARGUMENT, "lanes must be between 1 and 16");
    if (c.max_batch > 4096) Fail(VQM_INVALID_ARGUMENT, "max_batch must be between 1 and 4096");
    if (c.blur > kBlurRecursive) Fail(VQM_INVALID_ARGUMENT, "blur must be sampled (0) or recursive (1)");
    if (c.graph > 1) Fail(VQM_INVALID_ARGUMENT, "graph must be 0 or 1");
    ToMapForm(c.map_form, static_cast<PlaneFormat>(c.plane_format));
    // Validates strip_rows and strip_cols.
    BuildPlan(8, 8, static_cast<PlaneFormat>(c.plane_format), c.strip_rows, c.strip_cols);
    return c;
}

vqm_ss2_server_config WithDefaults(vqm_ss2_server_config c) {
    if (c.lanes == 0) c.lanes = 3;
    if (c.max_batch == 0) c.max_batch = 64;
    return c;
}

}  // namespace

Server::Server(const vqm_ss2_server_config& config)
    : config_(WithDefaults(ValidatedServer(config))),
      ctx_(DeviceContext::Get()),
      format_(static_cast<PlaneFormat>(config.plane_format)),
      map_(ToMapForm(config.map_form, static_cast<PlaneFormat>(config.plane_format))),
      pools_(static_cast<PlaneFormat>(config.plane_format), config.strip_rows, config.strip_cols,
             config.hbm_budget) {
    for (uint32_t i = 0; i < config_.lanes; ++i) {
        lanes_.push_back(std::make_unique<Lane>(config_.max_batch, config_.graph != 0));
        free_lanes_.push_back(config_.lanes - 1 - i);
    }
    dispatcher_ = std::thread([this] { DispatchLoop(); });
    completer_ = std::thread([this] { CompleteLoop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lock(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (dispatcher_.joinable()) dispatcher_.join();
    if (completer_.joinable()) completer_.join();
}

uint64_t Server::Enqueue(RequestPtr r) {
    r->ticket = NextId();
    r->submit_ns = NowNs();
    {
        std::lock_guard<std::mutex> lock(mu_);
        if (stop_) Fail(VQM_LOGIC_ERROR, "the server is shutting down");
        requests_[r->ticket] = r;
        queue_.push_back(r);
        ++stats_.submitted;
    }
    cv_.notify_all();
    return r->ticket;
}

uint64_t Server::SubmitDevice(const vqm_ss2_device_pair& pair) {
    PlanScales(pair.width, pair.height);
    if (pair.original == nullptr || pair.distorted == nullptr)
        Fail(VQM_INVALID_ARGUMENT, "device image data is null");
    auto r = std::make_shared<Request>();
    r->width = pair.width;
    r->height = pair.height;
    r->src[0] = pair.original;
    r->src[1] = pair.distorted;
    return Enqueue(std::move(r));
}

uint64_t Server::SubmitHost(const vqm_image_desc& a, const void* da, const vqm_image_desc& b,
                            const void* db) {
    PlanScales(a.width, a.height);
    CheckImage(a, a.width, a.height, "original");
    CheckImage(b, a.width, a.height, "distorted");
    if (!IsSrgb8(a.format) || !IsSrgb8(b.format))
        Fail(VQM_INVALID_ARGUMENT, "the server scores 8-bit full-range sRGB RGB images");
    if (da == nullptr || db == nullptr) Fail(VQM_INVALID_ARGUMENT, "image data is null");
    auto r = std::make_shared<Request>();
    r->width = a.width;
    r->height = a.height;
    r->host = true;
    r->src[0] = da;
    r->src[1] = db;
    r->image_bytes = RequiredBytes(Srgb8Format(), a.width, a.height);
    return Enqueue(std::move(r));
}

// Called with mu_ held; the request's workspace has already been returned.
void Server::FailRequest(const RequestPtr& r, int status, const std::string& message) {
    r->status = status;
    r->message = message;
    r->done = true;
    r->done_ns = NowNs();
    ++stats_.completed;
    ++stats_.failed;
}

void Server::DispatchLoop() {
    std::unique_lock<std::mutex> lock(mu_);
    for (;;) {
        cv_.wait(lock, [&] { return (stop_ && queue_.empty()) || (!queue_.empty() && !free_lanes_.empty()); });
        if (queue_.empty()) break;  // stopping, nothing left to launch
        const uint32_t lane = free_lanes_.back();
        free_lanes_.pop_back();

        std::vector<RequestPtr> batch;
        bool blocked = false;
        uint64_t seen = completions_;
        while (batch.size() < config_.max_batch && !queue_.empty()) {
            RequestPtr r = queue_.front();
            const bool alone = in_flight_.empty() && batch.empty();
            seen = completions_;
            lock.unlock();
            std::unique_ptr<Workspace> ws;
            int status = VQM_OK;
            std::string message;
            try {
                std::lock_guard<std::mutex> pl(pool_mu_);
                ws = pools_.Acquire(r->width, r->height, r->host);
                if (!ws && alone) {
                    status = VQM_INVALID_ARGUMENT;
                    message = "request of " + std::to_string(r->width) + "x" + std::to_string(r->height) +
                              " needs " + std::to_string(pools_.Need(r->width, r->height, r->host)) +
                              " workspace bytes, over the HBM budget of " + std::to_string(pools_.budget());
                }
            } catch (const Error& e) {
                status = e.status;
                message = e.message;
            } catch (const std::exception& e) {
                status = VQM_RUNTIME_ERROR;
                message = e.what();
            }
            lock.lock();
            if (status != VQM_OK) {
                queue_.pop_front();
                FailRequest(r, status, message);
                cv_.notify_all();
                continue;
            }
            if (!ws) {
                blocked = true;  // waits for workspaces held by batches in flight
                break;
            }
            queue_.pop_front();
            r->ws = std::move(ws);
            batch.push_back(std::move(r));
        }
        if (batch.empty()) {
            free_lanes_.push_back(lane);
            if (blocked) cv_.wait(lock, [&] { return completions_ != seen || in_flight_.empty(); });
            continue;
        }

        // Launch outside the lock: submitters and the completion thread go on.
        const uint32_t n = static_cast<uint32_t>(batch.size());
        std::vector<LaunchEntry> entries(n);
        lock.unlock();
        int status = VQM_OK;
        std::string message;
        const uint64_t t = NowNs();
        try {
            {
                std::lock_guard<std::mutex> pl(pool_mu_);
                for (uint32_t i = 0; i < n; ++i) entries[i].plan = &pools_.PlanFor(batch[i]->width, batch[i]->height);
            }
            for (uint32_t i = 0; i < n; ++i) {
                const Request& r = *batch[i];
                LaunchEntry& e = entries[i];
                e.ws = r.ws.get();
                for (int k = 0; k < 2; ++k) {
                    if (r.host)
                        e.host[k] = r.src[k];
                    else
                        e.src[k] = static_cast<const uint8_t*>(r.src[k]);
                }
                e.image_bytes = r.image_bytes;
            }
            lanes_[lane]->Launch(format_, map_, ctx_.lut, entries.data(), n, 1, config_.blur);
        } catch (const Error& e) {
            status = e.status;
            message = e.message;
        } catch (const std::exception& e) {
            status = VQM_RUNTIME_ERROR;
            message = e.what();
        }
        if (status != VQM_OK) {
            // Whatever was queued before the failure must end before the
            // workspaces are reused.
            lanes_[lane]->Synchronize();
            lanes_[lane]->Abandon();
            {
                std::lock_guard<std::mutex> pl(pool_mu_);
                for (RequestPtr& r : batch) pools_.Return(r->width, r->height, r->host, std::move(r->ws));
            }
            lock.lock();
            for (RequestPtr& r : batch) FailRequest(r, status, message);
            free_lanes_.push_back(lane);
            ++completions_;
            cv_.notify_all();
            continue;
        }
        lock.lock();
        for (RequestPtr& r : batch) {
            r->launch_ns = t;
            r->batch_size = n;
            r->lane = lane;
        }
        in_flight_.push_back(InFlight{lane, std::move(batch)});
        ++stats_.batches;
        stats_.batched_requests += n;
        stats_.max_batch = std::max<uint64_t>(stats_.max_batch, n);
        stats_.max_lanes_busy = std::max<uint32_t>(stats_.max_lanes_busy, static_cast<uint32_t>(in_flight_.size()));
        cv_.notify_all();
    }
    dispatcher_done_ = true;
    cv_.notify_all();
}

void Server::CompleteLoop() {
    std::unique_lock<std::mutex> lock(mu_);
    for (;;) {
        cv_.wait(lock, [&] { return !in_flight_.empty() || dispatcher_done_; });
        if (in_flight_.empty()) break;  // the dispatcher has stopped and everything finished
        std::vector<uint32_t> lanes;
        for (const InFlight& f : in_flight_) lanes.push_back(f.lane);
        lock.unlock();

        // Poll every lane in flight; batches on different lanes may finish in
        // any order.
        int found = -1;
        int status = VQM_OK;
        std::string message;
        for (size_t k = 0; k < lanes.size() && found < 0; ++k) {
            try {
                if (lanes_[lanes[k]]->Done()) found = static_cast<int>(k);
            } catch (const Error& e) {
                found = static_cast<int>(k);
                status = e.status;
                message = e.message;
            }
        }
        if (found < 0) {
            std::this_thread::yield();
            lock.lock();
            continue;
        }
        const uint64_t now = NowNs();
        const uint32_t lane = lanes[found];
        lock.lock();
        auto it = std::find_if(in_flight_.begin(), in_flight_.end(),
                               [&](const InFlight& f) { return f.lane == lane; });
        InFlight done = std::move(*it);
        in_flight_.erase(it);
        lock.unlock();

        if (status == VQM_OK) {
            try {
                lanes_[lane]->Wait();
            } catch (const Error& e) {
                status = e.status;
                message = e.message;
            }
        }
        if (status != VQM_OK) {
            // The batch's work may still touch its workspaces: whatever was
            // queued ends before they go back to the pools.
            lanes_[lane]->Synchronize();
            lanes_[lane]->Abandon();
        }
        for (uint32_t i = 0; i < done.requests.size(); ++i) {
            Request& r = *done.requests[i];
            if (status == VQM_OK) {
                const double* src = lanes_[lane]->Norms(i);
                r.norms.assign(src, src + kNorms);
                r.scales = PlanScales(r.width, r.height).count;
            }
        }
        {
            std::lock_guard<std::mutex> pl(pool_mu_);
            for (RequestPtr& r : done.requests) pools_.Return(r->width, r->height, r->host, std::move(r->ws));
        }
        lock.lock();
        for (RequestPtr& r : done.requests) {
            if (status != VQM_OK) {
                FailRequest(r, status, message);
                continue;
            }
            r->done = true;
            r->done_ns = now;
            ++stats_.completed;
        }
        free_lanes_.push_back(lane);
        ++completions_;
        cv_.notify_all();
    }
}

int Server::Wait(uint64_t ticket, vqm_ss2_request_result* out, std::string* message) {
    std::unique_lock<std::mutex> lock(mu_);
    auto it = requests_.find(ticket);
    if (ticket == 0 || it == requests_.end())
        Fail(VQM_INVALID_ARGUMENT, "unknown or already collected request ticket");
    RequestPtr r = it->second;
    cv_.wait(lock, [&] { return r->done; });
    requests_.erase(ticket);
    lock.unlock();
    if (out) {
        std::memset(out, 0, sizeof(*out));
        out->status = r->status;
        out->scales = r->scales;
        out->batch_size = r->batch_size;
        out->lane = r->lane;
        out->submit_ns = r->submit_ns;
        out->launch_ns = r->launch_ns;
        out->done_ns = r->done_ns;
        if (!r->norms.empty()) std::memcpy(out->norms, r->norms.data(), kNorms * sizeof(double));
    }
    if (message) *message = r->message;
    return r->status;
}

vqm_ss2_server_stats Server::Stats() {
    vqm_ss2_server_stats s;
    {
        std::lock_guard<std::mutex> lock(mu_);
        s = stats_;
        s.queued = static_cast<uint32_t>(queue_.size());
    }
    std::lock_guard<std::mutex> pl(pool_mu_);
    s.pool_bytes = pools_.bytes();
    s.peak_pool_bytes = pools_.peak();
    return s;
}

// ---- open-loop replay ---------------------------------------------------------

int Replay(Server& server, const vqm_ss2_replay_request* requests, uint32_t count,
           vqm_ss2_request_result* results, std::string* message) {
    std::vector<uint64_t> tickets(count, 0);
    std::vector<int> submit_status(count, VQM_OK);
    std::vector<std::string> submit_message(count);
    std::vector<std::string> wait_message(count);
    std::mutex m;
    std::condition_variable cv;
    uint32_t submitted = 0;
    const uint64_t t0 = NowNs() + 1000000;  // 1 ms to start the collector

    std::thread collector([&] {
        for (uint32_t i = 0; i < count; ++i) {
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return submitted > i; });
            }
            vqm_ss2_request_result& out = results[i];
            if (submit_status[i] != VQM_OK) {
                std::memset(&out, 0, sizeof(out));
                out.status = submit_status[i];
            } else {
                try {
                    server.Wait(tickets[i], &out, &wait_message[i]);
                } catch (const Error& e) {
                    std::memset(&out, 0, sizeof(out));
                    out.status = e.status;
                    wait_message[i] = e.message;
                }
            }
            out.submit_ns = t0 + requests[i].arrival_ns;
        }
    });

    for (uint32_t i = 0; i < count; ++i) {
        const vqm_ss2_replay_request& q = requests[i];
        const uint64_t due = t0 + q.arrival_ns;
        while (NowNs() < due) {
        }
        try {
            if (q.host) {
                vqm_image_desc d{};
                d.width = q.width;
                d.height = q.height;
                d.format = Srgb8Format();
                d.bytes = RequiredBytes(d.format, q.width, q.height);
                tickets[i] = server.SubmitHost(d, q.original, d, q.distorted);
            } else {
                vqm_ss2_device_pair p{q.original, q.distorted, q.width, q.height};
                tickets[i] = server.SubmitDevice(p);
            }
        } catch (const Error& e) {
            submit_status[i] = e.status;
            submit_message[i] = e.message;
        }
        {
            std::lock_guard<std::mutex> lk(m);
            submitted = i + 1;
        }
        cv.notify_one();
    }
    collector.join();
    for (uint32_t i = 0; i < count; ++i) {
        if (results[i].status != VQM_OK) {
            if (message)
                *message = "request " + std::to_string(i) + ": " +
                           (submit_status[i] != VQM_OK ? submit_message[i] : wait_message[i]);
            return results[i].status;
        }
    }
    return VQM_OK;
}

}  // namespace vqm::ss2
