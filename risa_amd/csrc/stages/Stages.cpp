// Stage implementations over the C-ABI (see risa/Stages.h).
#include "risa/Stages.h"

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>

#include "risa_synth.h"

namespace risa {

void check(int rc, const char* what)
{
    if (rc != RISA_OK)
        throw std::runtime_error(std::string("CUDA/HIP error in ") + what + ": " + risa_last_error());
}

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

int config_int(const std::string& path, const char* key)
{
    int v = 0;
    check(risa_config_lookup_int(path.c_str(), key, &v), key);
    return v;
}

int config_int_or(const std::string& path, const char* key, int fallback)
{
    int v = 0;
    return risa_config_lookup_int(path.c_str(), key, &v) == RISA_OK ? v : fallback;
}

std::string config_string_or(const std::string& path, const char* key, const std::string& fallback)
{
    char buf[4096];
    return risa_config_lookup_string(path.c_str(), key, buf, sizeof(buf)) == RISA_OK ? std::string(buf) : fallback;
}

int device_count()
{
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (n < 1) throw std::runtime_error("no HIP device visible");
    return n;
}

int processors_per_device(const std::string& configFile)
{
    const int n = config_int_or(configFile, "processorsPerDevice", 1);
    return n < 1 ? 1 : n;
}

int frame_batch(const std::string& configFile)
{
    const int n = config_int_or(configFile, "frameBatch", 1);
    return n < 1 ? 1 : n;
}

risa_geom read_geometry(const std::string& configFile)
{
    risa_geom g{};
    check(risa_geom_from_config(configFile.c_str(), &g), "readConfig");
    return g;
}

static std::size_t fan_elems(const risa_geom& g)
{
    return (std::size_t)g.numberOfFanDetectors * g.numberOfFanProjections;
}
static std::size_t par_elems(const risa_geom& g)
{
    return (std::size_t)g.numberOfParallelDetectors * g.numberOfParallelProjections;
}
static std::size_t img_elems(const risa_geom& g) { return (std::size_t)g.numberOfPixels * g.numberOfPixels; }

// ------------------------------------------------------- reference frames --
std::vector<uint16_t> read_module_frames(const std::string& stem, const risa_geom& g, int frames)
{
    const int modules = g.numberOfDetectorModules, D = g.numberOfFanDetectors, P = g.numberOfFanProjections;
    const int planes = g.numberOfPlanes;
    if (modules <= 0 || D % modules) throw std::runtime_error("numberOfDetectorModules must divide the detectors");
    const int per = D / modules;
    std::vector<uint16_t> out((std::size_t)frames * planes * P * D);
    for (int m = 0; m < modules; m++) {
        const std::string path = stem + std::to_string(m + 1) + ".fx";
        std::ifstream in(path, std::ios::binary);
        if (!in) throw std::runtime_error("Attenuation: reference file " + path + " could not be loaded");
        std::vector<uint16_t> content((std::size_t)frames * planes * P * per);
        in.read(reinterpret_cast<char*>(content.data()), (std::streamsize)(content.size() * sizeof(uint16_t)));
        if (!in) throw std::runtime_error("Attenuation: reference file " + path + " is too short");
        for (int b = 0; b < frames * planes; b++)
            for (int p = 0; p < P; p++) {
                const uint16_t* src = content.data() + ((std::size_t)b * P + p) * per;
                uint16_t* dst = out.data() + ((std::size_t)b * P + p) * D + (std::size_t)m * per;
                std::copy(src, src + per, dst);
            }
    }
    return out;
}

void write_module_frames(const std::string& stem, const risa_geom& g, int frames, const std::vector<uint16_t>& data)
{
    const int modules = g.numberOfDetectorModules, D = g.numberOfFanDetectors, P = g.numberOfFanProjections;
    const int planes = g.numberOfPlanes, per = D / modules;
    for (int m = 0; m < modules; m++) {
        std::vector<uint16_t> content((std::size_t)frames * planes * P * per);
        for (int b = 0; b < frames * planes; b++)
            for (int p = 0; p < P; p++) {
                const uint16_t* src = data.data() + ((std::size_t)b * P + p) * D + (std::size_t)m * per;
                std::copy(src, src + per, content.data() + ((std::size_t)b * P + p) * per);
            }
        std::ofstream out(stem + std::to_string(m + 1) + ".fx", std::ios::binary);
        out.write(reinterpret_cast<const char*>(content.data()), (std::streamsize)(content.size() * sizeof(uint16_t)));
        if (!out) throw std::runtime_error("could not write module file");
    }
}

std::vector<uint16_t> reference_frames(const std::string& configFile, const risa_geom& g, int R)
{
    std::string dir = config_string_or(configFile, "referenceInputPath", "");
    if (!dir.empty()) {
        if (dir.back() != '/') dir.push_back('/');
        return read_module_frames(dir + "ref_empty_tomograph_repaired_DetModNr_", g, R);
    }
    risa_synth_opts o;
    risa_synth_default_opts(&o);
    const uint64_t seed = (uint64_t)config_int_or(configFile, "synthReferenceSeed", 901);
    std::vector<uint16_t> out((std::size_t)R * g.numberOfPlanes * fan_elems(g));
    for (int b = 0; b < R * g.numberOfPlanes; b++)
        check(risa_synth_reference_frame(&g, seed * 1000003ull + (uint64_t)b + 1, &o, out.data() + b * fan_elems(g)),
              "risa_synth_reference_frame");
    return out;
}

// ------------------------------------------------------------------ H2D --
H2D::H2D(const std::string& configFile)
{
    const risa_geom g = read_geometry(configFile);
    const int pool = config_int(configFile, "memPoolSize_H2D");
    const std::size_t n = fan_elems(g);
    route_by_index_ = true;
    start("H2D", device_count(), processors_per_device(configFile), pool, n,
          [n](int, int, input_type& in, output_type& out, hipStream_t s) {
        if (in.size() < n) throw std::runtime_error("H2D: frame smaller than the fan sinogram");
        hip_check(hipMemcpyAsync(out.data(), in.data(), n * sizeof(uint16_t), hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync");
    });
}
H2D::~H2D()
{
    stop();
    std::fprintf(stderr, "H2D: WorstCaseTime: %gs; BestCaseTime: %gs;\n", loss_.worst_case_seconds(),
                 loss_.best_case_seconds());
    std::fprintf(stderr, "H2D: Could not reconstruct %lld elements; %g%% loss\n", loss_.lost(), loss_.loss_percent());
}

void H2D::process(input_type&& img)
{
    if (img.valid()) loss_.arrived(img.index());
    DeviceStage::process(std::move(img));
}

// ----------------------------------------------------------- Reordering --
Reordering::Reordering(const std::string& configFile)
{
    const risa_geom g = read_geometry(configFile);
    const int pool = config_int(configFile, "memPoolSize_Reordering");
    const int P = g.numberOfFanProjections, D = g.numberOfFanDetectors;
    start("Reordering", device_count(), processors_per_device(configFile), pool, fan_elems(g),
          [P, D](int, int, input_type& in, output_type& out, hipStream_t s) {
              check(risa_hip_reorder_run(in.data(), 1, P, D, out.data(), s), "Reordering::processor");
          });
}
Reordering::~Reordering() { stop(); }

// ---------------------------------------------------------- Attenuation --
// Device memory released at scope exit (set-up only).
template <class T>
struct SetupBuffer {
    T* p = nullptr;
    explicit SetupBuffer(std::size_t n) { hip_check(hipMalloc((void**)&p, n * sizeof(T)), "hipMalloc"); }
    ~SetupBuffer()
    {
        if (p) (void)hipFree(p);
    }
    SetupBuffer(const SetupBuffer&) = delete;
    SetupBuffer& operator=(const SetupBuffer&) = delete;
};

// Attenuation::init (Attenuation.cu:177-216) with the reference frames' defect
// repair and average done on each device (risa_hip_attenuation_prepare_reference).
Attenuation::Attenuation(const std::string& configFile)
{
    g_ = read_geometry(configFile);
    const int R = config_int(configFile, "numberOfReferenceFrames");
    const std::vector<uint16_t> frames = reference_frames(configFile, g_, R);
    const int devices = device_count(), lanes = processors_per_device(configFile);
    const int pool = config_int(configFile, "memPoolSize_attenuation");
    const std::size_t nref = (std::size_t)g_.numberOfPlanes * fan_elems(g_);
    const std::size_t ndark = (std::size_t)g_.numberOfPlanes * g_.numberOfFanDetectors;
    try {
        for (int d = 0; d < devices; d++) {
            glados::hip::ScopedDevice sd(d);
            SetupBuffer<uint16_t> dframes(frames.size());
            SetupBuffer<float> dref(nref), ddark(ndark);
            hip_check(hipMemcpy(dframes.p, frames.data(), frames.size() * sizeof(uint16_t), hipMemcpyHostToDevice),
                      "hipMemcpy");
            check(risa_hip_attenuation_prepare_reference(&g_, dframes.p, R, dref.p, ddark.p, nullptr),
                  "Attenuation::init");
            for (int l = 0; l < lanes; l++) {
                risa_attenuation_t h = nullptr;
                check(risa_hip_attenuation_create_device(&g_, dref.p, ddark.p, d, nullptr, &h), "Attenuation");
                h_.push_back(h);
            }
        }
    } catch (...) {
        for (auto h : h_) risa_hip_attenuation_destroy(h);
        h_.clear();
        throw;
    }
    start_processors(configFile, pool, devices, lanes);
}

Attenuation::Attenuation(const std::string& configFile, const std::vector<float>& ref, const std::vector<float>& dark)
{
    g_ = read_geometry(configFile);
    if (ref.size() != (std::size_t)g_.numberOfPlanes * fan_elems(g_) ||
        dark.size() != (std::size_t)g_.numberOfPlanes * g_.numberOfFanDetectors)
        throw std::runtime_error("Attenuation: reference/dark averages have the wrong size");
    init(configFile, ref, dark);
}

void Attenuation::init(const std::string& configFile, const std::vector<float>& ref, const std::vector<float>& dark)
{
    const int pool = config_int(configFile, "memPoolSize_attenuation");
    const int devices = device_count(), lanes = processors_per_device(configFile);
    for (int w = 0; w < devices * lanes; w++) {
        risa_attenuation_t h = nullptr;
        check(risa_hip_attenuation_create(&g_, ref.data(), dark.data(), w / lanes, &h), "Attenuation");
        h_.push_back(h);
    }
    start_processors(configFile, pool, devices, lanes);
}

void Attenuation::start_processors(const std::string& configFile, int pool, int devices, int lanes)
{
    start_batched("Attenuation", devices, lanes, pool, fan_elems(g_), frame_batch(configFile),
                  [this](int w, int, input_type* ins, output_type* outs, int n, hipStream_t s) {
                      std::vector<const uint16_t*> raw((std::size_t)n);
                      std::vector<float*> out((std::size_t)n);
                      std::vector<int32_t> planes((std::size_t)n);
                      for (int i = 0; i < n; i++) {
                          raw[(std::size_t)i] = ins[i].data();
                          out[(std::size_t)i] = outs[i].data();
                          planes[(std::size_t)i] = (int32_t)ins[i].plane();
                      }
                      check(risa_hip_attenuation_run_list(h_[(std::size_t)w], raw.data(), planes.data(), n,
                                                          out.data(), s),
                            "Attenuation::processor");
                  });
}

Attenuation::~Attenuation()
{
    stop();
    for (auto h : h_) risa_hip_attenuation_destroy(h);
}

// ------------------------------------------------------------- Fan2Para --
Fan2Para::Fan2Para(const std::string& configFile)
{
    g_ = read_geometry(configFile);
    const int pool = config_int(configFile, "memPoolSize_fan2Para");
    const int devices = device_count(), lanes = processors_per_device(configFile);
    // the rebinning tables are read-only: one handle per device serves its lanes
    for (int d = 0; d < devices; d++) {
        risa_fan2para_t h = nullptr;
        check(risa_hip_fan2para_create(&g_, d, &h), "Fan2Para");
        h_.push_back(h);
    }
    start_batched("Fan2Para", devices, lanes, pool, par_elems(g_), frame_batch(configFile),
                  [this, lanes](int w, int, input_type* ins, output_type* outs, int n, hipStream_t s) {
                      std::vector<const float*> fan((std::size_t)n);
                      std::vector<float*> par((std::size_t)n);
                      std::vector<int32_t> planes((std::size_t)n);
                      for (int i = 0; i < n; i++) {
                          fan[(std::size_t)i] = ins[i].data();
                          par[(std::size_t)i] = outs[i].data();
                          planes[(std::size_t)i] = (int32_t)ins[i].plane();
                      }
                      check(risa_hip_fan2para_run_list(h_[(std::size_t)(w / lanes)], fan.data(), planes.data(), n,
                                                       par.data(), s),
                            "Fan2Para::processor");
                  });
}

Fan2Para::~Fan2Para()
{
    stop();
    for (auto h : h_) risa_hip_fan2para_destroy(h);
}

// --------------------------------------------------------------- Filter --
Filter::Filter(const std::string& configFile)
{
    g_ = read_geometry(configFile);
    const int devices = device_count(), lanes = processors_per_device(configFile);
    for (int w = 0; w < devices * lanes; w++) {
        risa_filter_t h = nullptr;
        check(risa_hip_filter_create(&g_, w / lanes, RISA_FILTER_MODE_FFT32, &h), "Filter");
        h_.push_back(h);
    }
    start_batched("Filter", devices, lanes, 0, 0, frame_batch(configFile),
                  [this](int w, int, input_type* ins, output_type*, int n, hipStream_t s) {
                      std::vector<float*> sino((std::size_t)n);
                      for (int i = 0; i < n; i++) sino[(std::size_t)i] = ins[i].data();
                      check(risa_hip_filter_run_list(h_[(std::size_t)w], sino.data(), n, s), "Filter::processor");
                  });
}

Filter::~Filter()
{
    stop();
    for (auto h : h_) risa_hip_filter_destroy(h);
}

// ----------------------------------------------------- mixed precision --
bool mixed_precision(const std::string& configFile)
{
    const std::string p = config_string_or(configFile, "precision", "fp32");
    if (p == "fp32") return false;
    if (p == "mixed") return true;
    throw std::runtime_error("precision must be \"fp32\" or \"mixed\", not \"" + p + "\"");
}

FilterMixed::FilterMixed(const std::string& configFile)
{
    g_ = read_geometry(configFile);
    const int pool = config_int_or(configFile, "memPoolSize_filter", 16);
    const int devices = device_count(), lanes = processors_per_device(configFile);
    for (int w = 0; w < devices * lanes; w++) {
        risa_filter_t h = nullptr;
        check(risa_hip_filter_create(&g_, w / lanes, RISA_FILTER_MODE_GEMM_SPLIT, &h), "FilterMixed");
        h_.push_back(h);
    }
    start("FilterMixed", devices, lanes, pool, par_elems(g_),
          [this](int w, int, input_type& in, output_type& out, hipStream_t s) {
              check(risa_hip_filter_run_f16(h_[(std::size_t)w], in.data(), out.data(), 1, kMixedSinoScale, s),
                    "FilterMixed::processor");
          });
}

FilterMixed::~FilterMixed()
{
    stop();
    for (auto h : h_) risa_hip_filter_destroy(h);
}

BackprojectionMixed::BackprojectionMixed(const std::string& configFile)
{
    g_ = read_geometry(configFile);
    const int pool = config_int(configFile, "memPoolSize_backProjection");
    const int devices = device_count();
    const int lanes = std::max(1, config_int_or(configFile, "backProjectionLanes", processors_per_device(configFile)));
    // frames per kernel launch: the FP16 backprojector shares each pixel's
    // weights across the 4 slices of its sample layout, so it takes up to
    // `backProjectionBatch` queued frames at once (optional key, default 4)
    batch_ = std::max(1, std::min(config_int_or(configFile, "backProjectionBatch", 4), std::max(1, pool)));
    const std::size_t plane = (std::size_t)g_.numberOfParallelProjections * g_.numberOfParallelDetectors;
    for (int w = 0; w < devices * lanes; w++) {
        risa_bp_t h = nullptr;
        check(risa_hip_bp_create(&g_, w / lanes, &h), "BackprojectionMixed");
        h_.push_back(h);
        glados::hip::ScopedDevice sd(w / lanes);
        int variant = 0;
        if (w % lanes == 0) check(risa_hip_bp_autotune_f16(h, batch_, nullptr), "BackprojectionMixed autotune");
        check(risa_hip_bp_get_variant(h_[(std::size_t)(w - w % lanes)], &variant), "BackprojectionMixed variant");
        check(risa_hip_bp_set_variant(h, variant), "BackprojectionMixed variant");
        // per processor: contiguous staging for a batch of sinograms and images
        uint16_t* in16 = nullptr;
        float* out32 = nullptr;
        if (batch_ > 1) {
            hip_check(hipMalloc(&in16, (std::size_t)batch_ * plane * sizeof(uint16_t)), "BackprojectionMixed staging");
            hip_check(hipMalloc(&out32, (std::size_t)batch_ * img_elems(g_) * sizeof(float)),
                      "BackprojectionMixed staging");
        }
        in16_.push_back(in16);
        out32_.push_back(out32);
    }
    start_batched("BackprojectionMixed", devices, lanes, pool, img_elems(g_), batch_,
                  [this, plane](int w, int, input_type* ins, output_type* outs, int n, hipStream_t s) {
                      const risa_bp_t h = h_[(std::size_t)w];
                      if (n == 1) {
                          check(risa_hip_bp_run_f16(h, ins[0].data(), kMixedSinoScale, 1, outs[0].data(), s),
                                "BackprojectionMixed::processor");
                          return;
                      }
                      uint16_t* in16 = in16_[(std::size_t)w];
                      float* out32 = out32_[(std::size_t)w];
                      const std::size_t img = img_elems(g_);
                      for (int i = 0; i < n; i++)
                          hip_check(hipMemcpyAsync(in16 + (std::size_t)i * plane, ins[i].data(),
                                                   plane * sizeof(uint16_t), hipMemcpyDeviceToDevice, s),
                                    "BackprojectionMixed gather");
                      check(risa_hip_bp_run_f16(h, in16, kMixedSinoScale, n, out32, s),
                            "BackprojectionMixed::processor");
                      for (int i = 0; i < n; i++)
                          hip_check(hipMemcpyAsync(outs[i].data(), out32 + (std::size_t)i * img, img * sizeof(float),
                                                   hipMemcpyDeviceToDevice, s),
                                    "BackprojectionMixed scatter");
                  });
}

BackprojectionMixed::~BackprojectionMixed()
{
    stop();
    for (auto h : h_) risa_hip_bp_destroy(h);
    for (auto p : in16_)
        if (p) (void)hipFree(p);
    for (auto p : out32_)
        if (p) (void)hipFree(p);
}

// ------------------------------------------------------- Backprojection --
Backprojection::Backprojection(const std::string& configFile)
{
    g_ = read_geometry(configFile);
    const int pool = config_int(configFile, "memPoolSize_backProjection");
    // processor threads: the optional key backProjectionLanes, default
    // processorsPerDevice
    const int devices = device_count();
    const int lanes = std::max(1, config_int_or(configFile, "backProjectionLanes", processors_per_device(configFile)));
    // frames per kernel launch: up to `backProjectionBatch` queued frames
    // (optional key, default 4) go to one batched launch, whose tiles may
    // reconstruct two slices per workgroup; the tile shape is timed at the
    // frames the device's lanes then keep in flight
    batch_ = std::max(1, std::min(config_int_or(configFile, "backProjectionBatch", 4), std::max(1, pool)));
    for (int w = 0; w < devices * lanes; w++) {
        risa_bp_t h = nullptr;
        check(risa_hip_bp_create(&g_, w / lanes, &h), "Backprojection");
        h_.push_back(h);
        glados::hip::ScopedDevice sd(w / lanes);
        int variant = 0;
        if (w % lanes == 0) {
            // the optional key backProjectionVariant fixes the tile shape
            // (risa_hip_bp_variant_shape's numbering) instead of timing them
            const int fixed = config_int_or(configFile, "backProjectionVariant", -1);
            if (fixed >= 0) check(risa_hip_bp_set_variant(h, fixed), "Backprojection variant");
            else if (batch_ > 1) check(risa_hip_bp_autotune_batch(h, lanes * batch_, nullptr), "Backprojection autotune");
            else check(risa_hip_bp_autotune_slices(h, lanes, nullptr), "Backprojection autotune");
        }
        check(risa_hip_bp_get_variant(h_[(std::size_t)(w - w % lanes)], &variant), "Backprojection variant");
        check(risa_hip_bp_set_variant(h, variant), "Backprojection variant");
    }
    start_batched("Backprojection", devices, lanes, pool, img_elems(g_), batch_,
                  [this](int w, int, input_type* ins, output_type* outs, int n, hipStream_t s) {
                      std::vector<const float*> sino((std::size_t)n);
                      std::vector<float*> img((std::size_t)n);
                      for (int i = 0; i < n; i++) {
                          sino[(std::size_t)i] = ins[i].data();
                          img[(std::size_t)i] = outs[i].data();
                      }
                      check(risa_hip_bp_run_list(h_[(std::size_t)w], sino.data(), n, img.data(), s),
                            "Backprojection::processor");
                  });
}

Backprojection::~Backprojection()
{
    stop();
    for (auto h : h_) risa_hip_bp_destroy(h);
}

// -------------------------------------------------------------- Masking --
Masking::Masking(const std::string& configFile)
{
    g_ = read_geometry(configFile);
    const int devices = device_count(), lanes = processors_per_device(configFile);
    for (int w = 0; w < devices * lanes; w++) {
        risa_mask_t h = nullptr;
        check(risa_hip_mask_create(&g_, w / lanes, &h), "Masking");
        h_.push_back(h);
    }
    start("Masking", devices, lanes, 0, 0, [this](int w, int, input_type& img, output_type&, hipStream_t s) {
        check(risa_hip_mask_run(h_[(std::size_t)w], img.data(), s), "Masking::processor");
    });
}

Masking::~Masking()
{
    stop();
    for (auto h : h_) risa_hip_mask_destroy(h);
}

// ------------------------------------------------------------------ D2H --
D2H::D2H(const std::string& configFile) : rate_((std::size_t)std::max(1, config_int_or(configFile, "rateInterval", 10000)))
{
    const risa_geom g = read_geometry(configFile);
    const int pool = config_int(configFile, "memPoolSize_D2H");
    const std::size_t n = img_elems(g);
    start("D2H", device_count(), processors_per_device(configFile), pool, n,
          [n](int, int, input_type& in, output_type& out, hipStream_t s) {
        hip_check(hipMemcpyAsync(out.data(), in.data(), n * sizeof(float), hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync");
    });
}
D2H::~D2H()
{
    stop();
    std::fprintf(stderr, "D2H: Reconstructed %g Images/s in average (%zu images, %g images/s over the run).\n",
                 rate_.reconstruction_rate(), rate_.images(), rate_.images_per_second());
}

void D2H::process(input_type&& img)
{
    if (img.valid()) rate_.arrived(img.index());
    DeviceStage::process(std::move(img));
}

// ------------------------------------------------------- Reconstruction --
Reconstruction::Reconstruction(const std::string& configFile, bool module_major)
    : rate_((std::size_t)std::max(1, config_int_or(configFile, "rateInterval", 10000)))
{
    g_ = read_geometry(configFile);
    const int R = config_int(configFile, "numberOfReferenceFrames");
    const std::vector<uint16_t> frames = reference_frames(configFile, g_, R);
    const int devices = device_count();
    const int lanes = std::max(1, config_int_or(configFile, "reconstructionLanes", 3));
    batch_ = std::max(1, config_int_or(configFile, "reconstructionBatch", 32));
    // every lane's batch and the slices on their way to the sink
    const int pool = std::max(config_int(configFile, "memPoolSize_D2H"), 2 * lanes * batch_);
    const int mode = mixed_precision(configFile) ? RISA_CHAIN_MIXED : RISA_CHAIN_PLAIN;
    const std::size_t nf = fan_elems(g_), ni = img_elems(g_);
    std::vector<float> ref((std::size_t)g_.numberOfPlanes * nf), dark((std::size_t)g_.numberOfPlanes * g_.numberOfFanDetectors);
    route_by_index_ = true;
    try {
        for (int d = 0; d < devices; d++) {
            glados::hip::ScopedDevice sd(d);
            {
                // the Attenuation stage's reference average, read back for the chain
                SetupBuffer<uint16_t> dframes(frames.size());
                SetupBuffer<float> dref(ref.size()), ddark(dark.size());
                hip_check(hipMemcpy(dframes.p, frames.data(), frames.size() * sizeof(uint16_t), hipMemcpyHostToDevice),
                          "hipMemcpy");
                check(risa_hip_attenuation_prepare_reference(&g_, dframes.p, R, dref.p, ddark.p, nullptr),
                      "Reconstruction::init");
                hip_check(hipMemcpy(ref.data(), dref.p, ref.size() * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
                hip_check(hipMemcpy(dark.data(), ddark.p, dark.size() * sizeof(float), hipMemcpyDeviceToHost),
                          "hipMemcpy");
            }
            for (int l = 0; l < lanes; l++) {
                risa_chain_t h = nullptr;
                check(risa_hip_chain_create(&g_, ref.data(), dark.data(), d, mode, batch_, &h), "Reconstruction");
                h_.push_back(h);
                dev_.push_back(d);
                raw_.push_back(nullptr);
                img_.push_back(nullptr);
                if (module_major)
                    check(risa_hip_chain_set_input_layout(h, RISA_LAYOUT_MODULE_MAJOR), "Reconstruction");
                hip_check(hipMalloc(&raw_.back(), (std::size_t)batch_ * nf * sizeof(uint16_t)), "hipMalloc");
                hip_check(hipMalloc(&img_.back(), (std::size_t)batch_ * ni * sizeof(float)), "hipMalloc");
            }
        }
    } catch (...) {
        release();
        throw;
    }
    start_batched("Reconstruction", devices, lanes, pool, ni, batch_,
                  [this, nf, ni](int w, int, input_type* ins, output_type* outs, int n, hipStream_t s) {
                      // Frames take staging slots, and slices output buffers, in
                      // host address order (the pair ins[i] -> outs[i] is all the
                      // caller keeps), so each run of adjacent pooled buffers
                      // crosses PCIe as one copy.
                      auto by_address = [](const auto& a, const auto& b) { return a.data() < b.data(); };
                      std::sort(ins, ins + n, by_address);
                      std::sort(outs, outs + n, by_address);
                      std::vector<int32_t> planes((std::size_t)n);
                      for (int i = 0; i < n; i++) {
                          if (ins[i].size() < nf) throw std::runtime_error("Reconstruction: frame smaller than the fan sinogram");
                          planes[(std::size_t)i] = (int32_t)ins[i].plane();
                      }
                      for (int i = 0, j; i < n; i = j) {
                          for (j = i + 1; j < n && ins[j].data() == ins[j - 1].data() + nf; j++) {}
                          hip_check(hipMemcpyAsync(raw_[(std::size_t)w] + (std::size_t)i * nf, ins[i].data(),
                                                   (std::size_t)(j - i) * nf * sizeof(uint16_t), hipMemcpyHostToDevice, s),
                                    "hipMemcpyAsync");
                      }
                      // eager launches: the plane pattern of a queued batch varies from call to call
                      check(risa_hip_chain_run(h_[(std::size_t)w], raw_[(std::size_t)w], planes.data(), n,
                                               img_[(std::size_t)w], 0, s),
                            "Reconstruction::processor");
                      for (int i = 0, j; i < n; i = j) {
                          for (j = i + 1; j < n && outs[j].data() == outs[j - 1].data() + ni; j++) {}
                          hip_check(hipMemcpyAsync(outs[i].data(), img_[(std::size_t)w] + (std::size_t)i * ni,
                                                   (std::size_t)(j - i) * ni * sizeof(float), hipMemcpyDeviceToHost, s),
                                    "hipMemcpyAsync");
                      }
                  });
}

void Reconstruction::release()
{
    for (std::size_t i = 0; i < h_.size(); i++) {
        glados::hip::ScopedDevice sd(dev_[i]);
        risa_hip_chain_destroy(h_[i]);
        if (raw_[i]) (void)hipFree(raw_[i]);
        if (img_[i]) (void)hipFree(img_[i]);
    }
    h_.clear();
    raw_.clear();
    img_.clear();
    dev_.clear();
}

Reconstruction::~Reconstruction()
{
    stop();
    release();
    std::fprintf(stderr, "Reconstruction: Could not reconstruct %lld elements; %g%% loss\n", loss_.lost(),
                 loss_.loss_percent());
    std::fprintf(stderr, "Reconstruction: Reconstructed %g Images/s in average (%zu images, %g images/s over the run).\n",
                 rate_.reconstruction_rate(), rate_.images(), rate_.images_per_second());
}

void Reconstruction::process(input_type&& img)
{
    if (img.valid()) loss_.arrived(img.index());
    DeviceStage::process(std::move(img));
}

Reconstruction::output_type Reconstruction::wait()
{
    output_type out = DeviceStage::wait();
    if (out.valid()) rate_.arrived(out.index());
    return out;
}

// ---------------------------------------------------------- OfflineSaver --
static std::string config_string(const std::string& path, const char* key)
{
    char buf[4096];
    check(risa_config_lookup_string(path.c_str(), key, buf, sizeof(buf)), key);
    return std::string(buf);
}

OfflineSaver::OfflineSaver(const std::string& configFile)
{
    N_ = config_int(configFile, "numberOfPixels");
    outputPath_ = config_string(configFile, "outputPath");
    fileName_ = config_string(configFile, "outputFileName");
    (void)config_int(configFile, "numberOfDataFrames");   // read, as the reference does, and unused
    planes_ = config_int(configFile, "numberOfPlanes");
    bufferSize_ = config_int(configFile, "outputBufferSize");
    const std::string mode = config_string(configFile, "mode");
    if (mode == "offline") online_ = false;
    else if (mode == "online") online_ = true;
    else throw std::runtime_error("OfflineSaver: requested mode \"" + mode + "\" not supported");
    if (N_ < 1 || planes_ < 1 || bufferSize_ < 1) throw std::runtime_error("OfflineSaver: bad output configuration");
    buffers_.resize((std::size_t)planes_);
    fileIndex_.assign((std::size_t)planes_, 0u);
    last_.assign((std::size_t)planes_, std::chrono::steady_clock::now());
}

OfflineSaver::~OfflineSaver()
{
    for (int k = 0; k < planes_; k++) {
        try {
            writeSequence(k);
        } catch (const std::exception& e) {
            std::fprintf(stderr, "OfflineSaver: %s\n", e.what());
        }
    }
}

void OfflineSaver::saveImage(glados::Image<manager_type> image, std::string)
{
    const int k = (int)image.plane();
    if (k < 0 || k >= planes_) throw std::runtime_error("OfflineSaver: image of an unknown plane");
    if (image.size() < (std::size_t)N_ * N_) throw std::runtime_error("OfflineSaver: image smaller than N x N");
    const double latency = image.duration();
    if (latency < minLatency_) minLatency_ = latency;
    if (latency > maxLatency_) maxLatency_ = latency;
    if (!online_) {
        buffers_[(std::size_t)k].emplace_back(image.data(), image.data() + (std::size_t)N_ * N_);
        if ((int)buffers_[(std::size_t)k].size() >= bufferSize_) writeSequence(k);
        return;
    }
    const auto now = std::chrono::steady_clock::now();
    if (std::chrono::duration<double>(now - last_[(std::size_t)k]).count() < 0.02) return;
    const std::string path = outputPath_ + fileName_ + "_plane" + std::to_string(k) + "_0.tif";
    check(risa_tiff_write_f32(path.c_str(), image.data(), 1, N_, N_, RISA_TIFF_CLASSIC_BE), "OfflineSaver");
    last_[(std::size_t)k] = std::chrono::steady_clock::now();
}

void OfflineSaver::writeSequence(int k)
{
    auto& buf = buffers_[(std::size_t)k];
    if (buf.empty()) return;
    const std::string path = outputPath_ + fileName_ + "_plane" + std::to_string(k) + "_" +
                             std::to_string(fileIndex_[(std::size_t)k]) + "seq.tif";
    std::vector<float> pages(buf.size() * (std::size_t)N_ * N_);
    for (std::size_t i = 0; i < buf.size(); i++) std::copy(buf[i].begin(), buf[i].end(), pages.begin() + i * N_ * N_);
    check(risa_tiff_write_f32(path.c_str(), pages.data(), (int)buf.size(), N_, N_, RISA_TIFF_BIGTIFF), "OfflineSaver");
    ++fileIndex_[(std::size_t)k];
    buf.clear();
}

}  // namespace risa
