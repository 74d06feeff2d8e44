// Shared machinery of the GPU stages: per-device processor threads, each
// with its own HIP stream (default priority) and input queue, an output
// memory pool per device for allocating stages, and the results queue that
// wait() drains.
//
// The protocol is the reference stages' (e.g. Backprojection.cu:91-208):
// process() hands a valid Image to the queue of its device; the invalid
// sentinel makes every processor thread finish its queue and exit, after
// which the sentinel is put on the results queue.  Each frame's launches are
// stream-ordered and the stream is synchronised before the frame is handed
// on, so an output Image is complete when wait() returns it.
//
// Errors: a C-ABI status is turned into std::runtime_error with the
// library's message (the CHECK format of the reference).  An error on a
// processor thread ends that thread, is kept, and is rethrown by wait().
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include <glados/Image.h>
#include <glados/Queue.h>
#include <glados/hip/Memory.h>

#include "risa_hip.h"

namespace risa {

// Throws std::runtime_error for a non-zero C-ABI status.
void check(int rc, const char* what);
// Throws std::runtime_error for a HIP error.
void hip_check(hipError_t e, const char* what);

// Stage keys: integer / string lookups that throw on a missing key, and the
// optional form with a default.
int config_int(const std::string& path, const char* key);
int config_int_or(const std::string& path, const char* key, int fallback);
std::string config_string_or(const std::string& path, const char* key, const std::string& fallback);

// Number of visible devices the stages spread frames over.
int device_count();

// Processor threads (streams) per device of a stage: the optional key
// processorsPerDevice, default 1 (the reference runs one per device).
int processors_per_device(const std::string& configFile);

// Frames a processor of the Attenuation, Fan2Para and Filter stages takes
// from its queue at once (those already queued, none waited for) and runs
// in one launch: the optional key frameBatch, default 1 (the reference's
// one frame per call).
int frame_batch(const std::string& configFile);

template <class InT, class OutT>
class DeviceStage {
public:
    using input_type = InT;
    using output_type = OutT;

    DeviceStage(const DeviceStage&) = delete;
    DeviceStage& operator=(const DeviceStage&) = delete;
    // Derived stages call stop() first (their handles must outlive the
    // processor threads); this only catches a stage that did not.
    virtual ~DeviceStage()
    {
        if (!workers_.empty()) stop();
    }

    void process(input_type&& img)
    {
        if (img.valid()) {
            // frames from the host are dealt round-robin by index (H2D.cu:92-93);
            // device frames stay on their device
            const int d = (int)((route_by_index_ ? (int)img.index() : img.device()) % (int)queues_.size());
            queues_[(size_t)d]->push(std::move(img));
            return;
        }
        finish();
    }

    output_type wait()
    {
        output_type out = results_.take();
        if (!out.valid()) rethrow();
        return out;
    }

protected:
    // fn(worker, device, in, out, stream): worker = device * lanes + lane
    using compute_fn = std::function<void(int worker, int dev, input_type& in, output_type& out, hipStream_t s)>;
    // batched form: n frames at once, ins[i] -> outs[i]
    using batch_fn = std::function<void(int worker, int dev, input_type* ins, output_type* outs, int n,
                                        hipStream_t s)>;

    DeviceStage() = default;

    // Starts `lanes` processor threads per device, each with its own stream
    // (lanes = 1 is the reference's one processor per device; more keep
    // several frames of the stage in flight).  With out_elements > 0 the
    // stage allocates: each device gets a pool of pool_size output buffers
    // and fn writes into a fresh one; with 0 the stage works in place and
    // forwards the input Image itself (input_type must equal output_type).
    void start(const char* name, int devices, int lanes, int pool_size, std::size_t out_elements, compute_fn fn)
    {
        name_ = name;
        fn_ = std::move(fn);
        allocating_ = out_elements > 0;
        lanes_ = lanes < 1 ? 1 : lanes;
        for (int d = 0; d < devices; d++) {
            auto q = std::unique_ptr<glados::Queue<input_type>>(new glados::Queue<input_type>);
            int pool = -1;
            if (allocating_)
                pool = glados::MemoryPool<typename output_type::manager_type>::instance()->registerStage(
                    (std::size_t)pool_size, out_elements, d);
            for (int l = 0; l < lanes_; l++) {
                auto w = std::unique_ptr<Worker>(new Worker);
                w->index = d * lanes_ + l;
                w->device = d;
                w->queue = q.get();
                w->pool = pool;
                glados::hip::ScopedDevice sd(d);
                hip_check(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
                workers_.push_back(std::move(w));
            }
            queues_.push_back(std::move(q));
            pools_.push_back(pool);
        }
        for (auto& w : workers_) w->thread = std::thread(&DeviceStage::loop, this, w.get());
    }

    // Like start(), but a processor takes a frame and then, without waiting,
    // up to max_batch - 1 more that are already queued, and hands them to fn
    // together (one launch for the batch).  Frames still leave in the order
    // they were taken.  In-place stages (out_elements 0) get the inputs as
    // the outputs.
    void start_batched(const char* name, int devices, int lanes, int pool_size, std::size_t out_elements,
                       int max_batch, batch_fn fn)
    {
        max_batch_ = std::max(1, out_elements > 0 ? std::min(max_batch, pool_size) : max_batch);
        bfn_ = std::move(fn);
        start(name, devices, lanes, pool_size, out_elements, nullptr);
    }

    // Joins the processor threads (the sentinel is sent if process() never
    // saw one) and releases streams and pools.  Call from the derived
    // destructor, before its handles go.
    void stop()
    {
        if (!stopped_) finish();
        for (auto& w : workers_)
            if (w->stream) (void)hipStreamDestroy(w->stream);
        for (int pool : pools_)
            if (allocating_ && pool >= 0)
                glados::MemoryPool<typename output_type::manager_type>::instance()->freeMemory(pool);
        workers_.clear();
        queues_.clear();
        pools_.clear();
    }

private:
    struct Worker {
        int index = 0;
        int device = 0;
        hipStream_t stream = nullptr;
        int pool = -1;
        glados::Queue<input_type>* queue = nullptr;   // shared by the device's lanes
        std::thread thread;
        bool got_sentinel = false;                    // this lane's end-of-stream item was taken
    };

    void finish()
    {
        std::lock_guard<std::mutex> lock(finish_m_);
        if (stopped_) return;
        for (auto& w : workers_) w->queue->push(input_type());   // one sentinel per lane
        for (auto& w : workers_)
            if (w->thread.joinable()) w->thread.join();
        stopped_ = true;
        results_.push(output_type());
    }

    void loop(Worker* w)
    {
        try {
            glados::hip::ScopedDevice sd(w->device);
            if (bfn_) {
                batched_loop(w);
                return;
            }
            for (;;) {
                input_type in = w->queue->take();
                if (!in.valid()) {
                    w->got_sentinel = true;
                    break;
                }
                if constexpr (std::is_same<input_type, output_type>::value) {
                    if (!allocating_) {
                        fn_(w->index, w->device, in, in, w->stream);
                        hip_check(hipStreamSynchronize(w->stream), name_);
                        results_.push(std::move(in));
                        continue;
                    }
                }
                output_type out =
                    glados::MemoryPool<typename output_type::manager_type>::instance()->requestMemory(w->pool);
                fn_(w->index, w->device, in, out, w->stream);
                hip_check(hipStreamSynchronize(w->stream), name_);
                out.setIdx(in.index());
                out.setPlane(in.plane());
                out.setStart(in.start());
                results_.push(std::move(out));
            }
        } catch (...) {
            std::lock_guard<std::mutex> lock(err_m_);
            if (!error_) error_ = std::current_exception();
            // drain so producers never block on a dead consumer
            while (!w->got_sentinel) {
                input_type in = w->queue->take();
                if (!in.valid()) w->got_sentinel = true;
            }
        }
    }

    void batched_loop(Worker* w)
    {
        auto* pool = glados::MemoryPool<typename output_type::manager_type>::instance();
        std::vector<input_type> ins;
        std::vector<output_type> outs;
        bool end = false;
        while (!end) {
            ins.clear();
            outs.clear();
            input_type first = w->queue->take();
            if (!first.valid()) {
                w->got_sentinel = true;
                break;
            }
            ins.push_back(std::move(first));
            while ((int)ins.size() < max_batch_) {
                input_type more;
                if (!w->queue->try_take(more)) break;
                if (!more.valid()) {   // this lane's sentinel: finish the batch, then stop
                    w->got_sentinel = end = true;
                    break;
                }
                ins.push_back(std::move(more));
            }
            if constexpr (std::is_same<input_type, output_type>::value) {
                if (!allocating_) {
                    bfn_(w->index, w->device, ins.data(), ins.data(), (int)ins.size(), w->stream);
                    hip_check(hipStreamSynchronize(w->stream), name_);
                    for (auto& in : ins) results_.push(std::move(in));
                    continue;
                }
            }
            for (std::size_t i = 0; i < ins.size(); i++) outs.push_back(pool->requestMemory(w->pool));
            bfn_(w->index, w->device, ins.data(), outs.data(), (int)ins.size(), w->stream);
            hip_check(hipStreamSynchronize(w->stream), name_);
            for (std::size_t i = 0; i < ins.size(); i++) {
                outs[i].setIdx(ins[i].index());
                outs[i].setPlane(ins[i].plane());
                outs[i].setStart(ins[i].start());
                results_.push(std::move(outs[i]));
            }
        }
    }

    void rethrow()
    {
        std::lock_guard<std::mutex> lock(err_m_);
        if (error_) std::rethrow_exception(error_);
    }

protected:
    bool route_by_index_ = false;

private:
    const char* name_ = "stage";
    compute_fn fn_;
    batch_fn bfn_;
    int max_batch_ = 1;
    int lanes_ = 1;
    bool allocating_ = false;
    bool stopped_ = false;
    std::mutex finish_m_, err_m_;
    std::exception_ptr error_;
    std::vector<std::unique_ptr<Worker>> workers_;
    std::vector<std::unique_ptr<glados::Queue<input_type>>> queues_;   // per device
    std::vector<int> pools_;                                           // per device
    glados::Queue<output_type> results_;
};

}  // namespace risa
