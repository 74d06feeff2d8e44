// CPU self-test of the pipeline framework and the stages' host side: no GPU
// is touched.  Checks the queue bound and order, pool backpressure and
// release, sentinel propagation through Source -> Stage -> Stage -> Sink,
// the module-major reference file layout, that stages reject a
// configuration with a missing key, and the copy stages' frame accounting
// (lost indices, reconstruction rate) on a CPU pipeline whose source skips
// indices.
//
//   glados_selftest <config>      exit 0 when every check passes
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include <glados/Image.h>
#include <glados/Memory.h>
#include <glados/Queue.h>
#include <glados/pipeline/Pipeline.h>
#include <glados/pipeline/SinkStage.h>
#include <glados/pipeline/SourceStage.h>
#include <glados/pipeline/Stage.h>

#include "risa/Stages.h"

namespace {

int failures = 0, checks = 0;

void expect(bool ok, const char* what)
{
    checks++;
    if (!ok) {
        failures++;
        std::fprintf(stderr, "FAIL: %s\n", what);
    }
}

using HostInt = glados::PlainHostMemoryManager<int>;
using IntImage = glados::Image<HostInt>;

void test_queue()
{
    glados::Queue<int> q(10);
    std::atomic<bool> done{false};
    std::size_t max_seen = 0;
    std::thread prod([&] {
        for (int i = 0; i < 25; i++) q.push(int(i));
        done = true;
    });
    std::this_thread::sleep_for(std::chrono::milliseconds(50));
    max_seen = q.size();
    expect(!done.load(), "queue producer blocks at the bound");
    expect(max_seen == 10, "queue holds at most its bound (10)");
    bool ordered = true;
    for (int i = 0; i < 25; i++) ordered &= q.take() == i;
    prod.join();
    expect(ordered, "queue is FIFO");
}

void test_pool()
{
    auto* pool = glados::MemoryPool<HostInt>::instance();
    const int id = pool->registerStage(2, 16);
    std::atomic<bool> got{false};
    {
        IntImage a = pool->requestMemory(id);
        IntImage b = pool->requestMemory(id);
        expect(a.valid() && b.valid() && a.size() == 16, "pool hands out valid buffers of the stage size");
        expect(pool->available(id) == 0, "pool empty after two requests");
        std::thread t([&] {
            IntImage c = pool->requestMemory(id);
            got = c.valid();
        });
        std::this_thread::sleep_for(std::chrono::milliseconds(50));
        expect(!got.load(), "request blocks on an empty pool");
        { IntImage moved = std::move(a); }     // returned on destruction
        t.join();
        expect(got.load(), "a returned buffer wakes the waiting request");
    }
    expect(pool->available(id) == 2, "all buffers back in the pool");
    pool->freeMemory(id);
    bool threw = false;
    try {
        (void)pool->requestMemory(id);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    expect(threw, "request after freeMemory fails");
}

// CPU stages for the pipeline mechanics
class CountingLoader {
public:
    using manager_type = HostInt;
    explicit CountingLoader(int n) : n_(n) {}
    IntImage loadImage()
    {
        if (i_ >= n_) return IntImage();
        IntImage img(4, (std::size_t)i_);
        for (int k = 0; k < 4; k++) img.data()[k] = i_;
        ++i_;
        return img;
    }

private:
    int n_, i_ = 0;
};

class AddOne {
public:
    using input_type = IntImage;
    using output_type = IntImage;
    void process(input_type&& img)
    {
        if (img.valid())
            for (std::size_t k = 0; k < img.size(); k++) img.data()[k] += 1;
        q_.push(std::move(img));
    }
    output_type wait() { return q_.take(); }

private:
    glados::Queue<IntImage> q_;
};

struct Collected {
    std::vector<int> idx, val;
};

class Collector {
public:
    using manager_type = HostInt;
    explicit Collector(Collected* c) : c_(c) {}
    void saveImage(IntImage&& img, const std::string&)
    {
        c_->idx.push_back((int)img.index());
        c_->val.push_back(img.data()[0]);
    }

private:
    Collected* c_;
};

void test_pipeline()
{
    using namespace glados::pipeline;
    Collected got;
    Pipeline p;
    auto src = p.create<SourceStage<CountingLoader>>(37);
    auto s1 = p.create<Stage<AddOne>>();
    auto s2 = p.create<Stage<AddOne>>();
    auto sink = p.create<SinkStage<Collector>>("", "x", &got);
    p.connect(src, s1);
    p.connect(s1, s2);
    p.connect(s2, sink);
    p.run(src, s1, s2, sink);
    p.wait();
    bool ok = got.idx.size() == 37;
    for (std::size_t i = 0; ok && i < got.idx.size(); i++) ok = got.idx[i] == (int)i && got.val[i] == (int)i + 2;
    expect(ok, "37 frames pass two stages in order and the sentinel ends every stage");
}

// Source whose indices skip every k-th value (frames lost before the GPU);
// the stage in front counts them with H2D's counter, the one behind keeps
// D2H's rate meter.
class GappyLoader {
public:
    using manager_type = HostInt;
    GappyLoader(int n, int k) : n_(n), k_(k) {}
    IntImage loadImage()
    {
        if (emitted_ >= n_) return IntImage();
        if (k_ > 0 && next_ % k_ == k_ - 1) ++next_;
        IntImage img(4, (std::size_t)next_);
        ++next_;
        ++emitted_;
        return img;
    }

private:
    int n_, k_, next_ = 0, emitted_ = 0;
};

class CountLoss {
public:
    using input_type = IntImage;
    using output_type = IntImage;
    void process(input_type&& img)
    {
        if (img.valid()) loss.arrived(img.index());
        q_.push(std::move(img));
    }
    output_type wait() { return q_.take(); }
    risa::FrameLossCounter loss;

private:
    glados::Queue<IntImage> q_;
};

class MeterRate {
public:
    using input_type = IntImage;
    using output_type = IntImage;
    MeterRate() : rate(8) {}
    void process(input_type&& img)
    {
        if (img.valid()) {
            rate.arrived(img.index());
            std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
        q_.push(std::move(img));
    }
    output_type wait() { return q_.take(); }
    risa::RateMeter rate;

private:
    glados::Queue<IntImage> q_;
};

void test_accounting()
{
    using namespace glados::pipeline;
    Collected got;
    Pipeline p;
    // 50 frames, every 5th index left out: indices 0..61 minus 4, 9, ..., 59
    auto src = p.create<SourceStage<GappyLoader>>(50, 5);
    auto loss = p.create<Stage<CountLoss>>();
    auto rate = p.create<Stage<MeterRate>>();
    auto sink = p.create<SinkStage<Collector>>("", "x", &got);
    p.connect(src, loss);
    p.connect(loss, rate);
    p.connect(rate, sink);
    p.run(src, loss, rate, sink);
    p.wait();
    expect(got.idx.size() == 50, "gappy source: 50 frames reach the sink");
    expect(loss->loss.lost() == 12, "lost-frame count = the 12 skipped indices");
    expect(loss->loss.last_index() == 61, "last index seen is 61");
    const double pct = loss->loss.loss_percent();
    expect(pct > 19.67 && pct < 19.68, "loss percentage = lost / last index * 100 (12/61)");
    expect(loss->loss.worst_case_seconds() >= loss->loss.best_case_seconds(), "inter-arrival extremes ordered");
    expect(rate->rate.images() == 50 && rate->rate.windows() == 6, "rate meter: 6 full windows of 8 images");
    const double r = rate->rate.reconstruction_rate(), all = rate->rate.images_per_second();
    expect(r > 0 && r < 5000.0 + 1e-9 && all > 0 && all < 5000.0 + 1e-9, "rates bounded by the 200 us per image");
    // a gap-free stream from 0 loses nothing; out-of-order arrivals subtract
    risa::FrameLossCounter c;
    for (std::size_t i = 0; i < 10; i++) c.arrived(i);
    expect(c.lost() == 0 && c.loss_percent() == 0.0, "indices 0..9 lose nothing");
    c.arrived(15);
    c.arrived(13);
    expect(c.lost() == (15 - 9 - 1) + (13 - 15 - 1), "gap adds, out-of-order arrival subtracts (reference arithmetic)");
}

void test_module_files(const std::string& config)
{
    risa_geom g = risa::read_geometry(config);
    g.numberOfFanProjections = 5;          // keep the files small
    const int frames = 2;
    std::vector<uint16_t> data((std::size_t)frames * g.numberOfPlanes * g.numberOfFanProjections *
                               g.numberOfFanDetectors);
    for (std::size_t i = 0; i < data.size(); i++) data[i] = (uint16_t)(i * 2654435761u >> 16);
    char tmpl[] = "/tmp/risa_selftest_XXXXXX";
    const char* dir = mkdtemp(tmpl);
    expect(dir != nullptr, "temporary directory");
    if (!dir) return;
    const std::string stem = std::string(dir) + "/ref_empty_tomograph_repaired_DetModNr_";
    risa::write_module_frames(stem, g, frames, data);
    const auto back = risa::read_module_frames(stem, g, frames);
    expect(back == data, "module-major reference files re-sort to frame-major order");
    // a module's file holds its D/modules columns for every frame and row
    const int per = g.numberOfFanDetectors / g.numberOfDetectorModules;
    FILE* f = std::fopen((stem + "2.fx").c_str(), "rb");
    uint16_t first = 0;
    expect(f && std::fread(&first, 2, 1, f) == 1 && first == data[(std::size_t)per], "module 2 starts at column D/modules");
    if (f) std::fclose(f);
    for (int m = 1; m <= g.numberOfDetectorModules; m++) std::remove((stem + std::to_string(m) + ".fx").c_str());
    std::remove(dir);
}

void test_stage_config_errors(const std::string& config)
{
    // the stage keys are read before any device is touched
    std::FILE* in = std::fopen(config.c_str(), "rb");
    std::string text;
    char buf[4096];
    std::size_t n;
    while (in && (n = std::fread(buf, 1, sizeof(buf), in)) > 0) text.append(buf, n);
    if (in) std::fclose(in);
    const std::string key = "memPoolSize_H2D";
    const auto pos = text.find(key);
    expect(pos != std::string::npos, "config has memPoolSize_H2D");
    if (pos == std::string::npos) return;
    text.erase(pos, text.find('\n', pos) - pos);
    const std::string path = "/tmp/risa_selftest_nokey.cfg";
    std::FILE* out = std::fopen(path.c_str(), "wb");
    std::fwrite(text.data(), 1, text.size(), out);
    std::fclose(out);
    bool threw = false;
    try {
        risa::H2D h(path);
    } catch (const std::runtime_error& e) {
        threw = std::string(e.what()).find(key) != std::string::npos;
    }
    expect(threw, "H2D rejects a configuration without memPoolSize_H2D, naming the key");
    threw = false;
    try {
        risa::Backprojection b("/nonexistent/risa.cfg");
    } catch (const std::runtime_error&) {
        threw = true;
    }
    expect(threw, "a missing configuration file throws");
    std::remove(path.c_str());
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <config>\n", argv[0]);
        return 2;
    }
    test_queue();
    test_pool();
    test_pipeline();
    test_accounting();
    test_module_files(argv[1]);
    test_stage_config_errors(argv[1]);
    std::printf("glados_selftest: %d/%d checks passed\n", checks - failures, checks);
    return failures ? 1 : 0;
}
