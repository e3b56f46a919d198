// Request server with continuous batching over several HIP streams ("lanes").
//
// Submitters queue requests from any thread. The dispatcher thread forms a
// batch from the queue head whenever a lane is free: requests are admitted
// in arrival order while a workspace for their geometry can be taken from
// the pools within the HBM budget; the first request that does not fit ends
// the batch (it waits for workspaces to come back), unless nothing is in
// flight, in which case it can never fit and fails alone. The completion
// thread polls the lanes' events, copies each request's norms out of the
// lane's pinned read-back buffer, returns the workspaces and frees the lane.
// A request's arithmetic does not depend on its batch: every pair has its
// own planes, partials and norms, and the reduction order within a pair is
// fixed by its geometry.
#pragma once

#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ss2_runtime.h"

namespace vqm::ss2 {

uint64_t NowNs();

class Server {
public:
    explicit Server(const vqm_ss2_server_config& config);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    uint64_t SubmitDevice(const vqm_ss2_device_pair& pair);
    uint64_t SubmitHost(const vqm_image_desc& a, const void* da, const vqm_image_desc& b, const void* db);
    // Fills `out` and returns the request's status (0 when it succeeded).
    // Throws INVALID_ARGUMENT for an unknown ticket.
    int Wait(uint64_t ticket, vqm_ss2_request_result* out, std::string* message);
    vqm_ss2_server_stats Stats();

private:
    struct Request {
        uint64_t ticket = 0;
        uint32_t width = 0, height = 0;
        bool host = false;
        const void* src[2] = {nullptr, nullptr};
        uint64_t image_bytes = 0;
        uint64_t submit_ns = 0, launch_ns = 0, done_ns = 0;
        uint32_t batch_size = 0, lane = 0;
        bool done = false;
        int status = VQM_OK;
        std::string message;
        uint32_t scales = 0;
        std::unique_ptr<Workspace> ws;
        std::vector<double> norms;
    };
    using RequestPtr = std::shared_ptr<Request>;
    struct InFlight {
        uint32_t lane;
        std::vector<RequestPtr> requests;
    };

    uint64_t Enqueue(RequestPtr r);
    void DispatchLoop();
    void CompleteLoop();
    void FailRequest(const RequestPtr& r, int status, const std::string& message);

    vqm_ss2_server_config config_;
    DeviceContext& ctx_;
    PlaneFormat format_;
    MapForm map_;

    std::mutex mu_;                 // queue, requests, lanes, stats
    std::condition_variable cv_;    // any state change
    std::deque<RequestPtr> queue_;
    std::unordered_map<uint64_t, RequestPtr> requests_;
    std::vector<std::unique_ptr<Lane>> lanes_;
    std::vector<uint32_t> free_lanes_;
    std::deque<InFlight> in_flight_;
    uint64_t completions_ = 0;
    bool stop_ = false, dispatcher_done_ = false;
    vqm_ss2_server_stats stats_{};

    std::mutex pool_mu_;            // pools_
    WorkspacePools pools_;

    std::thread dispatcher_, completer_;
};

// Open-loop replay of a request trace against a server (vqm_ss2_server_replay).
int Replay(Server& server, const vqm_ss2_replay_request* requests, uint32_t count,
           vqm_ss2_request_result* results, std::string* message);

}  // namespace vqm::ss2
