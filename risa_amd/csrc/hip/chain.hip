// The whole per-frame chain on one device:
//   Attenuation -> Fan2Para -> Filter -> Backprojection (+ fused Masking)
// as RISA/example/main.cpp:95-134 wires the stages, without the per-stage
// host synchronisation of the reference (each stage there syncs its stream
// before handing the frame on, e.g. Backprojection.cu:203): launches are
// stream-ordered, and a captured hipGraph replays the launch sequence.
//
// PLAIN: frame by frame, FP32 throughout (FFT filter, FP32 backprojector).
//        Groups of `per` consecutive frames (up to 16; each stage runs once
//        per group) are dealt round-robin to `lanes` streams (own scratch
//        buffers each) forked from and joined back into the caller's stream,
//        so lanes x per frames are in flight -- the way the reference
//        overlaps frames across its per-stage threads.
// MIXED: groups of `per` consecutive frames (a multiple of 4, up to 32) dealt
//        to `lanes` streams the same way; per group the DFT-GEMM filter on
//        MFMA writes the FP16 sinogram in the 4-slice interleaved layout and
//        the FP16-sinogram / FP32-accumulate backprojection runs the group in
//        one launch, so one lane's front stages overlap the other's
//        backprojection.
#include "device_common.h"

#include <map>
#include <tuple>
#include <vector>

int risa_filter_run_oop(risa_filter_t h, const float* in, float* out, long rows, hipStream_t s);

namespace {
constexpr float kF16SinoScale = 256.0f;   // FP16 filtered sinogram stores 256 * value
constexpr int kPlainLanes = 2;            // streams in PLAIN mode
constexpr int kPlainMaxPer = 16;          // frames per lane step in PLAIN mode, at most
constexpr int kMixedLanes = 2;            // streams in MIXED mode
constexpr int kMixedMaxPer = 32;          // frames per lane step in MIXED mode, at most (a multiple of 4)
}  // namespace

// Graph executables a chain keeps (one per buffer set and call shape); the
// least recently replayed one is released beyond this.
constexpr size_t kMaxGraphs = 128;

struct ChainGraph {
    hipGraphExec_t exec = nullptr;
    unsigned long long last_use = 0;
};

struct risa_chain_s {
    int device = 0;
    int mode = 0;
    int max_frames = 0;
    int lanes = 1;
    int per = 1;   // frames per lane step
    risa_geom g{};
    risa_attenuation_t att = nullptr;
    risa_fan2para_t f2p = nullptr;
    risa_filter_t filt = nullptr;
    risa_bp_t bp = nullptr;
    risa::DeviceBuffer<float> att_buf, par_buf, filt_buf;   // `per` frame slots per lane
    risa::DeviceBuffer<uint16_t> f16_buf;                   // MIXED: `per` slots per lane, x4 layout
    std::vector<hipStream_t> lane_streams;
    std::vector<hipEvent_t> lane_events;   // [0] fork, [1..lanes] joins
    std::map<std::tuple<const void*, const void*, int, std::vector<int32_t>>, ChainGraph> graphs;
    unsigned long long uses = 0;

    // A graph executable, event or stream must not be released while work
    // that uses it may still run: hipGraphLaunch returns before the replay
    // ends, and the caller's stream is not known here, so the whole device
    // is drained first.
    void drain() const
    {
        int prev = -1;
        if (hipGetDevice(&prev) != hipSuccess) return;
        if (prev != device) (void)hipSetDevice(device);
        (void)hipDeviceSynchronize();
        if (prev != device) (void)hipSetDevice(prev);
    }
    void release_graphs()
    {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second.exec);
        graphs.clear();
    }
    ~risa_chain_s()
    {
        drain();
        release_graphs();
        for (hipEvent_t e : lane_events) (void)hipEventDestroy(e);
        for (hipStream_t s : lane_streams) (void)hipStreamDestroy(s);
        if (bp) risa_hip_bp_destroy(bp);
        if (filt) risa_hip_filter_destroy(filt);
        if (f2p) risa_hip_fan2para_destroy(f2p);
        if (att) risa_hip_attenuation_destroy(att);
    }
};

using namespace risa;

static void ok(int rc)
{
    if (rc != RISA_OK) fail(rc, "%s", risa_last_error());
}

extern "C" int risa_hip_chain_create(const risa_geom* g, const float* avg_ref, const float* avg_dark, int device,
                                     int mode, int max_frames, risa_chain_t* out)
{
    return guarded("risa_hip_chain_create", [&] {
        if (!g || !avg_ref || !avg_dark || !out || max_frames <= 0) fail(RISA_ERR_ARG, "bad argument");
        if (mode != RISA_CHAIN_PLAIN && mode != RISA_CHAIN_MIXED) fail(RISA_ERR_ARG, "unknown chain mode");
        DeviceGuard dg(device);
        auto h = new risa_chain_s();
        try {
            h->device = device;
            h->mode = mode;
            h->max_frames = max_frames;
            h->g = *g;
            ok(risa_hip_attenuation_create(g, avg_ref, avg_dark, device, &h->att));
            ok(risa_hip_fan2para_create(g, device, &h->f2p));
            ok(risa_hip_filter_create(g, device, mode == RISA_CHAIN_MIXED ? RISA_FILTER_MODE_GEMM_SPLIT : RISA_FILTER_MODE_FFT32,
                                      &h->filt));
            ok(risa_hip_bp_create(g, device, &h->bp));
            const size_t fan = (size_t)g->numberOfFanDetectors * g->numberOfFanProjections;
            const size_t par = (size_t)g->numberOfParallelDetectors * g->numberOfParallelProjections;
            if (mode == RISA_CHAIN_PLAIN) {
                h->lanes = std::min(kPlainLanes, max_frames);
                h->per = std::max(1, std::min(kPlainMaxPer, max_frames / h->lanes));
                const size_t slots = (size_t)h->lanes * h->per;
                h->att_buf.resize(fan * slots);
                h->par_buf.resize(par * slots);
                h->filt_buf.resize(par * slots);
            } else {
                // groups of a multiple of 4 frames: whole x4 slice groups per lane
                const int half = (max_frames + kMixedLanes - 1) / kMixedLanes;
                h->per = std::min(kMixedMaxPer, (half + 3) / 4 * 4);
                h->lanes = std::min(kMixedLanes, (max_frames + h->per - 1) / h->per);
                const size_t slots = (size_t)h->lanes * h->per;
                h->att_buf.resize(fan * slots);
                h->par_buf.resize(par * slots);
                h->f16_buf.resize(par * slots);   // x4 layout: per / 4 groups of 4 slices per lane
            }
            for (int k = 0; k < h->lanes; k++) {
                hipStream_t s;
                RISA_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
                h->lane_streams.push_back(s);
            }
            for (int k = 0; k <= h->lanes; k++) {
                hipEvent_t e;
                RISA_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                h->lane_events.push_back(e);
            }
            // backprojector tile shape, timed at the frames the chain keeps
            // in flight (plain: lanes x per slices, two-slice tiles included
            // when per > 1) or at a lane step's batch (mixed); timing the chain
            // itself per variant, eagerly, proved too noisy to rank the closest
            // shapes
            if (mode == RISA_CHAIN_PLAIN) {
                if (h->per > 1) ok(risa_hip_bp_autotune_batch(h->bp, h->lanes * h->per, nullptr));
                else ok(risa_hip_bp_autotune_slices(h->bp, h->lanes, nullptr));
            } else {
                ok(risa_hip_bp_autotune_f16(h->bp, h->per, nullptr));
            }
            RISA_HIP_CHECK(hipDeviceSynchronize());
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

static void enqueue(risa_chain_t h, const uint16_t* raw, const int32_t* planes, int frames, float* img, hipStream_t s)
{
    const risa_geom& g = h->g;
    const size_t fan = (size_t)g.numberOfFanDetectors * g.numberOfFanProjections;
    const size_t par = (size_t)g.numberOfParallelDetectors * g.numberOfParallelProjections;
    const size_t out = (size_t)g.numberOfPixels * g.numberOfPixels;
    const int per = h->per;
    const int groups = (frames + per - 1) / per;
    const int lanes = std::min(h->lanes, groups);
    RISA_HIP_CHECK(hipEventRecord(h->lane_events[0], s));
    for (int k = 0; k < lanes; k++) RISA_HIP_CHECK(hipStreamWaitEvent(h->lane_streams[k], h->lane_events[0], 0));
    for (int q = 0; q < groups; q++) {
        const int k = q % lanes, f = q * per, n = std::min(per, frames - f);
        hipStream_t ls = h->lane_streams[k];
        float* att = h->att_buf.ptr + (size_t)k * per * fan;
        float* pb = h->par_buf.ptr + (size_t)k * per * par;
        if (h->mode == RISA_CHAIN_PLAIN) {
            float* fb = h->filt_buf.ptr + (size_t)k * per * par;
            if (n == 1) {
                ok(risa_hip_attenuation_run(h->att, raw + f * fan, planes[f], att, ls));
                ok(risa_hip_fan2para_run(h->f2p, att, planes[f], pb, ls));
                ok(risa_filter_run_oop(h->filt, pb, fb, g.numberOfParallelProjections, ls));
                ok(risa_hip_bp_run(h->bp, fb, img + f * out, ls));
            } else {
                ok(risa_hip_attenuation_run_batch(h->att, raw + f * fan, planes + f, n, att, ls));
                ok(risa_hip_fan2para_run_batch(h->f2p, att, planes + f, n, pb, ls));
                ok(risa_filter_run_oop(h->filt, pb, fb, (long)n * g.numberOfParallelProjections, ls));
                ok(risa_hip_bp_run_batch(h->bp, fb, n, img + f * out, ls));
            }
        } else {
            uint16_t* hb = h->f16_buf.ptr + (size_t)k * per * par;
            ok(risa_hip_attenuation_run_batch(h->att, raw + f * fan, planes + f, n, att, ls));
            ok(risa_hip_fan2para_run_batch(h->f2p, att, planes + f, n, pb, ls));
            ok(risa_hip_filter_run_f16x4(h->filt, pb, hb, n, kF16SinoScale, ls));
            ok(risa_hip_bp_run_f16x4(h->bp, hb, kF16SinoScale, n, img + f * out, ls));
        }
    }
    for (int k = 0; k < lanes; k++) {
        RISA_HIP_CHECK(hipEventRecord(h->lane_events[k + 1], h->lane_streams[k]));
        RISA_HIP_CHECK(hipStreamWaitEvent(s, h->lane_events[k + 1], 0));
    }
}

extern "C" int risa_hip_chain_run(risa_chain_t h, const uint16_t* raw, const int32_t* planes, int frames, float* img,
                                  int use_graph, void* stream)
{
    return guarded("risa_hip_chain_run", [&] {
        if (!h || !raw || !planes || !img || frames < 0) fail(RISA_ERR_ARG, "bad argument");
        if (frames > h->max_frames) fail(RISA_ERR_ARG, "frames (%d) exceeds max_frames (%d)", frames, h->max_frames);
        if (frames == 0) return;
        for (int f = 0; f < frames; f++)
            if (planes[f] < 0 || planes[f] >= h->g.numberOfPlanes) fail(RISA_ERR_ARG, "plane out of range");
        hipStream_t s = as_stream(stream);
        if (!use_graph) {
            enqueue(h, raw, planes, frames, img, s);
            return;
        }
        if (!s) fail(RISA_ERR_ARG, "graph replay needs a non-null stream");
        auto key = std::make_tuple((const void*)raw, (const void*)img, frames, std::vector<int32_t>(planes, planes + frames));
        auto it = h->graphs.find(key);
        if (it == h->graphs.end()) {
            if (h->graphs.size() >= kMaxGraphs) {
                // release the least recently replayed executable, once no
                // replay of it can still be running
                auto lru = h->graphs.begin();
                for (auto jt = h->graphs.begin(); jt != h->graphs.end(); ++jt)
                    if (jt->second.last_use < lru->second.last_use) lru = jt;
                DeviceGuard dg(h->device);
                RISA_HIP_CHECK(hipDeviceSynchronize());
                RISA_HIP_CHECK(hipGraphExecDestroy(lru->second.exec));
                h->graphs.erase(lru);
            }
            hipGraph_t graph;
            RISA_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                enqueue(h, raw, planes, frames, img, s);
            } catch (...) {
                hipGraph_t dummy = nullptr;
                if (hipStreamEndCapture(s, &dummy) == hipSuccess && dummy) (void)hipGraphDestroy(dummy);
                throw;
            }
            RISA_HIP_CHECK(hipStreamEndCapture(s, &graph));
            hipGraphExec_t exec;
            const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            RISA_HIP_CHECK(ie);
            it = h->graphs.emplace(key, ChainGraph{exec, 0}).first;
        }
        it->second.last_use = ++h->uses;
        RISA_HIP_CHECK(hipGraphLaunch(it->second.exec, s));
    });
}

extern "C" int risa_hip_chain_intermediate(risa_chain_t h, int which, const void** ptr)
{
    return guarded("risa_hip_chain_intermediate", [&] {
        if (!h || !ptr) fail(RISA_ERR_ARG, "null argument");
        switch (which) {
            case 0: *ptr = h->att_buf.ptr; break;
            case 1: *ptr = h->par_buf.ptr; break;
            case 2: *ptr = h->mode == RISA_CHAIN_PLAIN ? (const void*)h->filt_buf.ptr : (const void*)h->f16_buf.ptr; break;
            default: fail(RISA_ERR_ARG, "which must be 0, 1 or 2");
        }
    });
}

// Captured graphs hold the launches of the layout and variant they were
// captured with: a change drains the device, releases them, and the next
// graph run captures anew.
extern "C" int risa_hip_chain_set_input_layout(risa_chain_t h, int layout)
{
    return guarded("risa_hip_chain_set_input_layout", [&] {
        if (!h) fail(RISA_ERR_ARG, "null handle");
        int cur = 0;
        ok(risa_hip_attenuation_get_input_layout(h->att, &cur));
        if (cur == layout) return;
        h->drain();
        h->release_graphs();
        ok(risa_hip_attenuation_set_input_layout(h->att, layout));
    });
}

extern "C" int risa_hip_chain_set_variant(risa_chain_t h, int variant)
{
    return guarded("risa_hip_chain_set_variant", [&] {
        if (!h) fail(RISA_ERR_ARG, "null handle");
        int cur = 0;
        ok(risa_hip_bp_get_variant(h->bp, &cur));
        if (cur == variant) return;
        h->drain();
        h->release_graphs();
        ok(risa_hip_bp_set_variant(h->bp, variant));
    });
}

/* Graph executables the chain holds (tooling; no reference counterpart). */
extern "C" int risa_hip_chain_graph_count(risa_chain_t h, int* count)
{
    return guarded("risa_hip_chain_graph_count", [&] {
        if (!h || !count) fail(RISA_ERR_ARG, "null argument");
        *count = (int)h->graphs.size();
    });
}

extern "C" int risa_hip_chain_get_variant(risa_chain_t h, int* variant)
{
    return guarded("risa_hip_chain_get_variant", [&] {
        if (!h || !variant) fail(RISA_ERR_ARG, "null argument");
        ok(risa_hip_bp_get_variant(h->bp, variant));
    });
}

extern "C" int risa_hip_chain_destroy(risa_chain_t h)
{
    return guarded("risa_hip_chain_destroy", [&] { delete h; });
}
