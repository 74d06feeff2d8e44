// The reconstruction pipeline as RISA/example/main.cpp:95-134 wires it:
//   Source -> H2D -> Attenuation -> Fan2Para -> Filter -> Backprojection
//          [-> Masking] -> D2H -> Sink
// with frames streamed host -> device -> host through the stage queues.
//
// The source replays seeded synthetic uint16 fan frames held in pinned host
// memory (the role of OfflineLoader_perfTest.cu:59-132, which replays
// recorded frames from RAM); the sink counts slices, measures latency and
// keeps the first few slices for checking.
//
//   risa_pipeline <config> [--frames N] [--unique U] [--save K FILE] [--masking] [--module-major] [--tiff]
//                 [--skip-every K] [--sync spin|yield|block] [--fused]
//
// --fused: Source -> Reconstruction -> Sink, the device stages and both copies
// as one stage that runs queued frames in batches through the chain (keys
// reconstructionLanes, reconstructionBatch); not with --masking.
//
// precision = "mixed" in the configuration wires FilterMixed and
// BackprojectionMixed (FP16 filtered sinograms) in place of Filter and
// Backprojection.
// --sync: how the stages' host threads wait for their streams (the device's
// scheduling flag, set before any other HIP call; default: the runtime's).
// --skip-every K: the source leaves out every index i with i % K == K - 1
// (frames lost before the GPU), which H2D's accounting counts.
//
// --module-major: the source holds ROFEX module-major frames (27 modules x
// P_f x 16, D_f = 432) and a Reordering stage follows H2D, as
// RISA/example/main.cpp:112-126 wires it.
// --tiff: the sink also hands every slice to risa::OfflineSaver, which writes
// TIFF files as the config's output keys say (outputPath, mode, ...).
//
// Prints one JSON line: frames, seconds, slices/s, mean and max latency, the
// precision, H2D's lost-frame count and loss percentage, and D2H's rates.
// --save writes K slices as records {int32 index, int32 frame, int32 plane,
// float img[N*N]} (frame = the seeded frame id, see below).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include <glados/pipeline/Pipeline.h>
#include <glados/pipeline/SinkStage.h>
#include <glados/pipeline/SourceStage.h>
#include <glados/pipeline/Stage.h>

#include "risa/Stages.h"
#include "risa_synth.h"

namespace {

// Frame id u of plane k is seeded as phantom 1000 + u, noise 55 + u*planes + k.
struct Provenance {
    std::mutex m;
    std::map<std::size_t, std::pair<int, int>> of_index;   // index -> (frame id, plane)
};

class SyntheticLoader {
public:
    using manager_type = glados::hip::HostMemoryManager<uint16_t>;

    SyntheticLoader(const std::string& configFile, long frames, int unique, Provenance* prov, bool module_major,
                    long skip_every)
        : frames_(frames), skip_every_(skip_every), prov_(prov)
    {
        g_ = risa::read_geometry(configFile);
        const int Pf = g_.numberOfFanProjections, Df = g_.numberOfFanDetectors;
        if (module_major && Df != 27 * 16)
            throw std::runtime_error("--module-major needs 27 modules of 16 detectors (numberOfFanDetectors = 432)");
        std::vector<uint16_t> frame;
        const int planes = g_.numberOfPlanes;
        const std::size_t n = (std::size_t)g_.numberOfFanDetectors * g_.numberOfFanProjections;
        auto* pool = glados::MemoryPool<manager_type>::instance();
        const int count = unique * planes;
        pool_ = pool->registerStage((std::size_t)count, n, 0);
        risa_synth_opts o;
        risa_synth_default_opts(&o);
        // every pooled buffer holds one distinct frame for good; handing a
        // buffer out replays the frame it holds
        std::vector<glados::Image<manager_type>> held;
        for (int b = 0; b < count; b++) {
            auto img = pool->requestMemory(pool_);
            const int u = b / planes, k = b % planes;
            risa::check(risa_synth_fan_frame(&g_, k, 1000ull + (uint64_t)u, 55ull + (uint64_t)(u * planes + k), &o,
                                             img.data()),
                        "risa_synth_fan_frame");
            if (module_major) {
                // [P_f][module * 16 + d] -> [module][P_f][16], what the detector sends
                frame.assign(img.data(), img.data() + n);
                for (int p = 0; p < Pf; p++)
                    for (int m = 0; m < 27; m++)
                        for (int d = 0; d < 16; d++)
                            img.data()[((std::size_t)m * Pf + p) * 16 + d] = frame[(std::size_t)p * Df + m * 16 + d];
            }
            ids_[img.data()] = {u, k};
            held.push_back(std::move(img));
        }
    }

    glados::Image<manager_type> loadImage()
    {
        if (emitted_ >= frames_) return glados::Image<manager_type>();
        if (skip_every_ > 0 && next_ % skip_every_ == skip_every_ - 1) ++next_;   // index lost upstream
        auto img = glados::MemoryPool<manager_type>::instance()->requestMemory(pool_);
        const auto id = ids_.at(img.data());
        img.setIdx((std::size_t)next_);
        img.setPlane((std::size_t)id.second);
        img.setStart(std::chrono::high_resolution_clock::now());
        {
            std::lock_guard<std::mutex> lock(prov_->m);
            prov_->of_index[(std::size_t)next_] = id;
        }
        ++next_;
        ++emitted_;
        return img;
    }

private:
    risa_geom g_{};
    long frames_ = 0, next_ = 0, emitted_ = 0, skip_every_ = 0;
    int pool_ = -1;
    Provenance* prov_;
    std::map<const uint16_t*, std::pair<int, int>> ids_;
};

struct SinkState {
    std::mutex m;
    long count = 0;
    double lat_sum = 0.0, lat_max = 0.0;
    std::chrono::high_resolution_clock::time_point last;
    int keep = 0;
    std::map<std::size_t, std::vector<float>> kept;
};

class StatsSaver {
public:
    using manager_type = glados::hip::HostMemoryManager<float>;

    StatsSaver(SinkState* st, const std::string& tiffConfig) : st_(st)
    {
        if (!tiffConfig.empty()) tiff_.reset(new risa::OfflineSaver(tiffConfig));
    }

    void saveImage(glados::Image<manager_type>&& img, const std::string& path)
    {
        const double lat = img.duration();
        {
            std::lock_guard<std::mutex> lock(st_->m);
            st_->count++;
            st_->lat_sum += lat;
            st_->lat_max = std::max(st_->lat_max, lat);
            st_->last = std::chrono::high_resolution_clock::now();
            if ((long)img.index() < st_->keep)
                st_->kept[img.index()] = std::vector<float>(img.data(), img.data() + img.size());
        }
        if (tiff_) tiff_->saveImage(std::move(img), path);
    }

private:
    SinkState* st_;
    std::unique_ptr<risa::OfflineSaver> tiff_;
};

struct Options {
    std::string config;
    long frames = 200;
    int unique = 8, keep = 0;
    long skip_every = 0;
    bool masking = false, module_major = false, tiff = false, fused = false;
    unsigned sync_flag = 0;   // hipDeviceSchedule*, 0: the runtime's default
    std::string save_path;
};

struct RunStats {
    double secs = 0.0;
    long long lost = 0;
    double loss_percent = 0.0, d2h_rate = 0.0, d2h_images_per_s = 0.0;
};

// The reference's wiring (main.cpp:95-134) with the given Filter and
// Backprojection stage types.
template <class FilterT, class BackprojectionT>
RunStats run_pipeline(const Options& o, Provenance& prov, SinkState& st)
{
    using namespace glados::pipeline;
    using Source = SourceStage<SyntheticLoader>;
    using Sink = SinkStage<StatsSaver>;
    Pipeline pipeline;
    auto source = pipeline.create<Source>(o.config, o.frames, o.unique, &prov, o.module_major, o.skip_every);
    auto h2d = pipeline.create<Stage<risa::H2D>>(o.config);
    auto reordering = o.module_major ? pipeline.create<Stage<risa::Reordering>>(o.config) : nullptr;
    auto attenuation = pipeline.create<Stage<risa::Attenuation>>(o.config);
    auto fan2para = pipeline.create<Stage<risa::Fan2Para>>(o.config);
    auto filter = pipeline.create<Stage<FilterT>>(o.config);
    auto backprojection = pipeline.create<Stage<BackprojectionT>>(o.config);
    auto masking_stage = o.masking ? pipeline.create<Stage<risa::Masking>>(o.config) : nullptr;
    auto d2h = pipeline.create<Stage<risa::D2H>>(o.config);
    auto sink = pipeline.create<Sink>("", "IMG", &st, o.tiff ? o.config : std::string());

    pipeline.connect(source, h2d);
    if (reordering) {
        pipeline.connect(h2d, reordering);
        pipeline.connect(reordering, attenuation);
    } else {
        pipeline.connect(h2d, attenuation);
    }
    pipeline.connect(attenuation, fan2para);
    pipeline.connect(fan2para, filter);
    pipeline.connect(filter, backprojection);
    if (masking_stage) {
        pipeline.connect(backprojection, masking_stage);
        pipeline.connect(masking_stage, d2h);
    } else {
        pipeline.connect(backprojection, d2h);
    }
    pipeline.connect(d2h, sink);

    const auto t0 = std::chrono::high_resolution_clock::now();
    pipeline.run(source, h2d);
    if (reordering) pipeline.run(reordering);
    pipeline.run(attenuation, fan2para, filter, backprojection);
    if (masking_stage) pipeline.run(masking_stage);
    pipeline.run(d2h, sink);
    pipeline.wait();
    const auto t1 = std::chrono::high_resolution_clock::now();
    RunStats r;
    r.secs = std::chrono::duration<double>(t1 - t0).count();
    r.lost = h2d->lostFrames();
    r.loss_percent = h2d->lossPercent();
    r.d2h_rate = d2h->reconstructionRate();
    r.d2h_images_per_s = d2h->imagesPerSecond();
    return r;
}

// Source -> Reconstruction (H2D .. D2H as one batched stage) -> Sink.
RunStats run_fused(const Options& o, Provenance& prov, SinkState& st)
{
    using namespace glados::pipeline;
    using Source = SourceStage<SyntheticLoader>;
    using Sink = SinkStage<StatsSaver>;
    Pipeline pipeline;
    auto source = pipeline.create<Source>(o.config, o.frames, o.unique, &prov, o.module_major, o.skip_every);
    auto rec = pipeline.create<Stage<risa::Reconstruction>>(o.config, o.module_major);
    auto sink = pipeline.create<Sink>("", "IMG", &st, o.tiff ? o.config : std::string());
    pipeline.connect(source, rec);
    pipeline.connect(rec, sink);
    const auto t0 = std::chrono::high_resolution_clock::now();
    pipeline.run(source, rec, sink);
    pipeline.wait();
    const auto t1 = std::chrono::high_resolution_clock::now();
    RunStats r;
    r.secs = std::chrono::duration<double>(t1 - t0).count();
    r.lost = rec->lostFrames();
    r.loss_percent = rec->lossPercent();
    r.d2h_rate = rec->reconstructionRate();
    r.d2h_images_per_s = rec->imagesPerSecond();
    return r;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr,
                     "usage: %s <config> [--frames N] [--unique U] [--save K FILE] [--masking] [--module-major] "
                     "[--tiff] [--skip-every K] [--sync spin|yield|block] [--fused]\n",
                     argv[0]);
        return 2;
    }
    Options o;
    o.config = argv[1];
    for (int i = 2; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--frames" && i + 1 < argc) o.frames = std::atol(argv[++i]);
        else if (a == "--unique" && i + 1 < argc) o.unique = std::atoi(argv[++i]);
        else if (a == "--save" && i + 2 < argc) {
            o.keep = std::atoi(argv[++i]);
            o.save_path = argv[++i];
        } else if (a == "--masking") o.masking = true;
        else if (a == "--module-major") o.module_major = true;
        else if (a == "--tiff") o.tiff = true;
        else if (a == "--fused") o.fused = true;
        else if (a == "--skip-every" && i + 1 < argc) o.skip_every = std::atol(argv[++i]);
        else if (a == "--sync" && i + 1 < argc) {
            const std::string m = argv[++i];
            if (m == "spin") o.sync_flag = hipDeviceScheduleSpin;
            else if (m == "yield") o.sync_flag = hipDeviceScheduleYield;
            else if (m == "block") o.sync_flag = hipDeviceScheduleBlockingSync;
            else {
                std::fprintf(stderr, "--sync takes spin, yield or block\n");
                return 2;
            }
        } else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    if (o.fused && o.masking) {
        std::fprintf(stderr, "--fused runs the fused mask only (no --masking)\n");
        return 2;
    }
    if (o.skip_every == 1) {
        std::fprintf(stderr, "--skip-every needs K >= 2\n");
        return 2;
    }

    if (o.sync_flag) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        for (int d = 0; d < n; d++)
            if (hipSetDevice(d) != hipSuccess || hipSetDeviceFlags(o.sync_flag) != hipSuccess) {
                std::fprintf(stderr, "cannot set the scheduling flag on device %d\n", d);
                return 1;
            }
        if (n > 0) (void)hipSetDevice(0);
    }
    Provenance prov;
    SinkState st;
    st.keep = o.keep;
    try {
        const bool mixed = risa::mixed_precision(o.config);
        const RunStats r = o.fused ? run_fused(o, prov, st)
                           : mixed ? run_pipeline<risa::FilterMixed, risa::BackprojectionMixed>(o, prov, st)
                                   : run_pipeline<risa::Filter, risa::Backprojection>(o, prov, st);
        if (!o.save_path.empty()) {
            std::FILE* f = std::fopen(o.save_path.c_str(), "wb");
            if (!f) throw std::runtime_error("cannot open " + o.save_path);
            for (auto& kv : st.kept) {
                const auto id = prov.of_index.at(kv.first);
                const int32_t hdr[3] = {(int32_t)kv.first, id.first, id.second};
                std::fwrite(hdr, sizeof(hdr), 1, f);
                std::fwrite(kv.second.data(), sizeof(float), kv.second.size(), f);
            }
            std::fclose(f);
        }
        std::printf("{\"frames\": %ld, \"seconds\": %.4f, \"slices_per_s\": %.1f, \"latency_ms_mean\": %.3f, "
                    "\"latency_ms_max\": %.3f, \"saved\": %zu, \"precision\": \"%s\", \"lost_frames\": %lld, "
                    "\"loss_percent\": %.4f, \"d2h_reconstruction_rate\": %.1f, \"d2h_images_per_s\": %.1f, "
                    "\"stages\": \"%s\"}\n",
                    st.count, r.secs, st.count / r.secs, st.count ? st.lat_sum / st.count : 0.0, st.lat_max,
                    st.kept.size(), mixed ? "mixed" : "fp32", r.lost, r.loss_percent, r.d2h_rate,
                    r.d2h_images_per_s, o.fused ? "fused" : "staged");
        if (st.count != o.frames) {
            std::fprintf(stderr, "expected %ld slices, got %ld\n", o.frames, st.count);
            return 1;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "risa_pipeline: %s\n", e.what());
        return 1;
    }
    return 0;
}
