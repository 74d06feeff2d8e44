// The reconstruction stages as glados pipeline stages.  Each keeps the
// reference class's shape -- input_type / output_type, a constructor taking
// the libconfig file, process(input_type&&), wait() -> output_type -- and
// reads the same keys; the work is done by the HIP kernels behind the C-ABI
// (include/risa_hip.h).
//
//   H2D            Copy/H2D.cu:84-169            host uint16 frame -> device
//   Reordering     Reordering.cu:33-204          ROFEX module-major -> detector-major
//   Attenuation    Attenuation.cu:48-175         uint16 fan sinogram -> float
//   Fan2Para       Fan2Para.cu:56-249            fan -> parallel sinogram
//   Filter         Filter.cu:54-143              in place
//   Backprojection Backprojection.cu:46-208      parallel sinogram -> slice
//   Masking        Masking.cu:46-169             in place
//   D2H            Copy/D2H.cu:83-148            device slice -> host
//   Reconstruction (no counterpart)              H2D .. D2H as one batched stage
//   OfflineSaver   OfflineSaver.cu:66-178        sink: TIFF sequences / latest image
#pragma once

#include <chrono>
#include <cstdint>
#include <string>
#include <vector>

#include <glados/Image.h>
#include <glados/hip/Memory.h>

#include "Accounting.h"
#include "DeviceStage.h"
#include "risa_hip.h"

namespace risa {

template <class T>
using DeviceImage = glados::Image<glados::hip::DeviceMemoryManager<T>>;
template <class T>
using HostImage = glados::Image<glados::hip::HostMemoryManager<T>>;

// Geometry and filter keys of a configuration file (throws on error).
risa_geom read_geometry(const std::string& configFile);

// The R reference frames per plane the Attenuation stage averages: read from
// the module files under referenceInputPath
// (<path>ref_empty_tomograph_repaired_DetModNr_<m>.fx, m = 1..modules, each
// module's columns for every frame, re-sorted as Attenuation.cu:302-321
// does), or, when referenceInputPath is empty, synthesised with the seeded
// generator (flat field Poisson(60000) + dark N(200, 5)).
std::vector<uint16_t> reference_frames(const std::string& configFile, const risa_geom& g, int R);

// The module-major file layout of the reference frames (Attenuation.cu:275-321):
// module m holds, for each frame f, plane k and projection p, its
// D/modules detector values.  Read (and the inverse, for tests).
std::vector<uint16_t> read_module_frames(const std::string& stem, const risa_geom& g, int frames);
void write_module_frames(const std::string& stem, const risa_geom& g, int frames, const std::vector<uint16_t>& data);

// Copies host frames to the devices, dealt round-robin by index
// (H2D.cu:84-169).  process() counts the indices missing between arrivals
// (H2D.cu:102-107); the destructor logs the count and the inter-arrival
// extremes as the reference does (:80-81).
class H2D : public DeviceStage<HostImage<uint16_t>, DeviceImage<uint16_t>> {
public:
    explicit H2D(const std::string& configFile);
    ~H2D();
    void process(input_type&& img);
    long long lostFrames() const { return loss_.lost(); }
    double lossPercent() const { return loss_.loss_percent(); }
    double bestCaseTime() const { return loss_.best_case_seconds(); }
    double worstCaseTime() const { return loss_.worst_case_seconds(); }

private:
    FrameLossCounter loss_;
};

// Permutes each ROFEX frame from its module-major layout (27 modules x P_f x
// 16) to [P_f][D_f] (risa_hip_reorder_run); keys numberOfFanDetectors,
// memPoolSize_Reordering, samplingRate, scanRate as Reordering::readConfig.
class Reordering : public DeviceStage<DeviceImage<uint16_t>, DeviceImage<uint16_t>> {
public:
    explicit Reordering(const std::string& configFile);
    ~Reordering();
};

class Attenuation : public DeviceStage<DeviceImage<uint16_t>, DeviceImage<float>> {
public:
    explicit Attenuation(const std::string& configFile);
    // With the averaged reference [planes][P_f][D_f] and dark [planes][D_f] given.
    Attenuation(const std::string& configFile, const std::vector<float>& avgReference,
                const std::vector<float>& avgDark);
    ~Attenuation();

private:
    void init(const std::string& configFile, const std::vector<float>& ref, const std::vector<float>& dark);
    void start_processors(const std::string& configFile, int pool, int devices, int lanes);
    risa_geom g_{};
    std::vector<risa_attenuation_t> h_;
};

class Fan2Para : public DeviceStage<DeviceImage<float>, DeviceImage<float>> {
public:
    explicit Fan2Para(const std::string& configFile);
    ~Fan2Para();

private:
    risa_geom g_{};
    std::vector<risa_fan2para_t> h_;
};

class Filter : public DeviceStage<DeviceImage<float>, DeviceImage<float>> {
public:
    explicit Filter(const std::string& configFile);
    ~Filter();

private:
    risa_geom g_{};
    std::vector<risa_filter_t> h_;
};

// The mixed-precision form of Filter (selected by precision = "mixed" in the
// configuration): the DFT-GEMM on the matrix cores (FP16 hi + lo split, FP32
// accumulation) writing each filtered sinogram as FP16 [P][D], value *
// kMixedSinoScale.  Allocating (pool memPoolSize_filter, default 16).
constexpr float kMixedSinoScale = 256.0f;
class FilterMixed : public DeviceStage<DeviceImage<float>, DeviceImage<uint16_t>> {
public:
    explicit FilterMixed(const std::string& configFile);
    ~FilterMixed();

private:
    risa_geom g_{};
    std::vector<risa_filter_t> h_;
};

// Backprojection of FilterMixed's FP16 sinograms with FP32 accumulation
// (risa_hip_bp_run_f16); masking fused as in Backprojection.
class BackprojectionMixed : public DeviceStage<DeviceImage<uint16_t>, DeviceImage<float>> {
public:
    explicit BackprojectionMixed(const std::string& configFile);
    ~BackprojectionMixed();

private:
    risa_geom g_{};
    int batch_ = 4;                  // frames per launch (key backProjectionBatch)
    std::vector<risa_bp_t> h_;
    std::vector<uint16_t*> in16_;    // per processor: [batch][P][D] staging
    std::vector<float*> out32_;      // per processor: [batch][N][N] staging
};

// precision = "mixed" in the configuration (default "fp32").
bool mixed_precision(const std::string& configFile);

// FP32 backprojection with the fused mask; up to backProjectionBatch (default
// 4) queued frames per launch (risa_hip_bp_run_list: no copies).
class Backprojection : public DeviceStage<DeviceImage<float>, DeviceImage<float>> {
public:
    explicit Backprojection(const std::string& configFile);
    ~Backprojection();

private:
    risa_geom g_{};
    int batch_ = 4;                // frames per launch (key backProjectionBatch)
    std::vector<risa_bp_t> h_;
};

class Masking : public DeviceStage<DeviceImage<float>, DeviceImage<float>> {
public:
    explicit Masking(const std::string& configFile);
    ~Masking();

private:
    risa_geom g_{};
    std::vector<risa_mask_t> h_;
};

// Copies slices back to the host (D2H.cu:83-148).  process() keeps the
// reference's reconstruction rate: a running mean of the rate over windows of
// rateInterval images (optional key, default 10000 as D2H.cu:85), logged at
// destruction (:42); imagesPerSecond() is the whole-run rate.
class D2H : public DeviceStage<DeviceImage<float>, HostImage<float>> {
public:
    explicit D2H(const std::string& configFile);
    ~D2H();
    void process(input_type&& img);
    double reconstructionRate() const { return rate_.reconstruction_rate(); }
    double imagesPerSecond() const { return rate_.images_per_second(); }
    std::size_t images() const { return rate_.images(); }

private:
    RateMeter rate_;
};

// The device part of the chain -- H2D, Attenuation, Fan2Para, Filter,
// Backprojection (mask fused), D2H -- as one stage, host frames in, host
// slices out.  No reference counterpart: it spans the six stages
// RISA/example/main.cpp:95-134 wires between loader and saver and replaces
// their per-frame queue hand-offs and stream synchronisations by one per
// batch.  A processor takes up to reconstructionBatch queued frames (default
// 32), copies them to its device staging, runs them as one chain call
// (risa_hip_chain_run: plain FP32, or mixed with precision = "mixed";
// module-major input when constructed so) and copies the slices out, all on
// its stream; reconstructionLanes processors per device (default 3) keep
// batches in flight, so one lane's copies overlap another's kernels.  A
// frame's host buffer is held until its batch is done: the source's pool
// must exceed lanes x batch frames per device.  H2D's lost-frame accounting
// (arrivals) and D2H's reconstruction rate (slices handed on) are kept.
class Reconstruction : public DeviceStage<HostImage<uint16_t>, HostImage<float>> {
public:
    explicit Reconstruction(const std::string& configFile, bool module_major = false);
    ~Reconstruction();
    void process(input_type&& img);
    output_type wait();
    long long lostFrames() const { return loss_.lost(); }
    double lossPercent() const { return loss_.loss_percent(); }
    double reconstructionRate() const { return rate_.reconstruction_rate(); }
    double imagesPerSecond() const { return rate_.images_per_second(); }

private:
    void release();
    risa_geom g_{};
    int batch_ = 32;
    std::vector<risa_chain_t> h_;   // per processor
    std::vector<uint16_t*> raw_;    // per processor: [batch][P_f][D_f] device staging
    std::vector<float*> img_;       // per processor: [batch][N][N] device staging
    std::vector<int> dev_;          // per processor: its device
    FrameLossCounter loss_;
    RateMeter rate_;
};

// Saver for glados::pipeline::SinkStage (OfflineSaver.cu:66-178), keys
// numberOfPixels, outputPath, outputFileName, numberOfDataFrames,
// numberOfPlanes, outputBufferSize and mode as OfflineSaver::readConfig.
// "offline": images are kept per plane and written as a multi-page BigTIFF
// <outputPath><outputFileName>_plane<k>_<n>seq.tif whenever outputBufferSize
// have gathered, the rest at destruction.  "online": at most every 20 ms per
// plane the newest image replaces <outputPath><outputFileName>_plane<k>_0.tif
// (big-endian classic TIFF).
class OfflineSaver {
public:
    using manager_type = glados::hip::HostMemoryManager<float>;
    explicit OfflineSaver(const std::string& configFile);
    ~OfflineSaver();
    void saveImage(glados::Image<manager_type> image, std::string path);
    double minLatency() const { return minLatency_; }
    double maxLatency() const { return maxLatency_; }

private:
    void writeSequence(int plane);
    int N_ = 0, planes_ = 0, bufferSize_ = 0;
    bool online_ = false;
    std::string outputPath_, fileName_;
    std::vector<std::vector<std::vector<float>>> buffers_;
    std::vector<unsigned> fileIndex_;
    std::vector<std::chrono::steady_clock::time_point> last_;
    double minLatency_ = 1e30, maxLatency_ = 0.0;
};

}  // namespace risa
