// Driver for the reference's own defect-detector functions
// (RISA/risaLib/src/DetectorInterpolation/interpolationFunctions.h:28-97),
// compiled against that header where it lies in the reference checkout
// (see oracle/Makefile, target `ref`).  Used only to write the golden
// vectors under tests/golden/; never shipped, never part of a GPU run.
//
// Input file : uint32 D, uint32 P, double threshMin, double threshMax,
//              uint16 data[P][D]
// Output file: int32 flags[D], uint16 repaired[P][D]
#include <cstdint>
#include <cstdio>
#include <vector>

#include "interpolationFunctions.h"

int main(int argc, char** argv)
{
    if (argc != 3) { std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]); return 2; }
    std::FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 3;
    uint32_t D = 0, P = 0;
    double tmin = 0, tmax = 0;
    if (std::fread(&D, 4, 1, in) != 1 || std::fread(&P, 4, 1, in) != 1 ||
        std::fread(&tmin, 8, 1, in) != 1 || std::fread(&tmax, 8, 1, in) != 1) return 4;
    std::vector<unsigned short> data((size_t)D * P);
    if (std::fread(data.data(), 2, data.size(), in) != data.size()) return 5;
    std::fclose(in);

    // The 17-tap window exactly as Attenuation::init builds it (Attenuation.cu:179-182).
    std::vector<double> w{0.5, 1.0, 1.0, 1.0, 1.5, 2.0, 3.0, 3.5, 2.0, 3.5, 3.0, 2.0, 1.5, 1.0, 1.0, 1.0, 0.5};
    double sum = 0.0;
    for (double v : w) sum += v;
    for (double& v : w) v = (1.0 / sum) * v;

    std::vector<int> flags((size_t)P * D, 0);   // sized as Attenuation.cu:192 does
    findDefectDetectors(data.data(), w, flags, D, P, tmin, tmax);
    interpolateDefectDetectors(data.data(), flags, D, P);

    std::FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 6;
    for (uint32_t d = 0; d < D; d++) { int32_t f = flags[d]; std::fwrite(&f, 4, 1, out); }
    std::fwrite(data.data(), 2, data.size(), out);
    std::fclose(out);
    return 0;
}
