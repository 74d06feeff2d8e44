/*
 * Oracle restatement of the Attenuation stage (TEST INFRASTRUCTURE ONLY).
 *
 * Reference: RISA/risaLib/src/Attenuation/Attenuation.cu and
 *            RISA/risaLib/src/DetectorInterpolation/interpolationFunctions.h.
 *
 * Built with -ffp-contract=off and no -march flags, so float expressions are
 * evaluated in IEEE single precision exactly as written (x86-64 SSE).
 */
#include "risa_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

size_t orc_geom_size(void) { return sizeof(orc_geom); }

/* double -> unsigned int as x86-64 gcc converts it (via a 64-bit signed
 * conversion), which is what `unsigned int y = std::round(...)` does. */
static unsigned int to_uint(double v) { return (unsigned int)(long long)v; }

/* Attenuation::relevantAreaMask, Attenuation.cu:327-381. */
void orc_area_mask(const orc_geom* g, float* mask)
{
    const unsigned int D = (unsigned int)g->numberOfFanDetectors;
    const int P = g->numberOfFanProjections;
    const unsigned int xa = (unsigned)g->xa, xb = (unsigned)g->xb, xc = (unsigned)g->xc;
    const unsigned int xd = (unsigned)g->xd, xe = (unsigned)g->xe, xf = (unsigned)g->xf;
    unsigned int ya, yb, yc, yd, ye, yMin, yMax;
    const double lowerLimit = (g->lowerLimOffset + g->sourceOffset) / 360.0;   /* :331 */
    const double upperLimit = (g->upperLimOffset + g->sourceOffset) / 360.0;   /* :332 */
    const size_t total = (size_t)P * D;
    for (size_t i = 0; i < total; i++) mask[i] = 1.0f;                          /* :335 */

    ya = to_uint(round(lowerLimit * P));                                        /* :337 */
    yb = ya;
    yc = to_uint(round(upperLimit * P));                                        /* :339 */
    yd = yc;
    const double m = ((double)ya - (double)yd) / ((double)xa - (double)xd);     /* :343 */
    ye = to_uint(round((double)yc + ((double)xe - (double)xc) * m));            /* :345 */

    for (unsigned int x = 0; x <= xa; x++) {                                    /* :347-352 */
        yMin = ya;
        yMax = to_uint(round(ye + m * x));
        for (unsigned int y = yMin; y < yMax; y++) mask[x + y * D] = 0.0f;
    }
    for (unsigned int x = xa; x <= xc; x++) {                                   /* :354-359 */
        yMin = to_uint(round(ya + m * (x - xa)));
        yMax = to_uint(round(ye + m * x));
        for (unsigned int y = yMin; y < yMax; y++) mask[x + y * D] = 0.0f;
    }
    for (unsigned int x = xc; x <= xd; x++) {                                   /* :361-366 */
        yMin = to_uint(round(ya + m * (x - xa)));
        yMax = yd;
        for (unsigned int y = yMin; y < yMax; y++) mask[x + y * D] = 0.0f;
    }
    for (unsigned int x = xb; x <= xf; x++) {                                   /* :368-373 */
        yMin = yb;
        yMax = to_uint(round(yb + m * (x - xb)));
        for (unsigned int y = yMin; y < yMax; y++) mask[x + y * D] = 0.0f;
    }
    /* :375-380 -- iterator + double truncates toward zero. */
    long lo = (long)(lowerLimit * (int)D * P);
    long hi = (long)(upperLimit * P * (int)D);
    if (lo > (long)total) lo = (long)total;
    for (long i = 0; i < lo; i++) mask[i] = 0.0f;
    if (hi < 0) hi = 0;
    for (long i = hi; i < (long)total; i++) mask[i] = 0.0f;
}

/* Attenuation::init, Attenuation.cu:179-182. */
void orc_defect_window(double* f)
{
    static const double w[17] = {0.5, 1.0, 1.0, 1.0, 1.5, 2.0, 3.0, 3.5, 2.0,
                                 3.5, 3.0, 2.0, 1.5, 1.0, 1.0, 1.0, 0.5};
    double sum = 0.0;
    for (int i = 0; i < 17; i++) sum += w[i];
    for (int i = 0; i < 17; i++) f[i] = (1.0 / sum) * w[i];
}

/* findDefectDetectors, interpolationFunctions.h:28-71. */
void orc_find_defects(const uint16_t* data, const double* ff, int* defect,
                      unsigned D, unsigned P, double threshMin, double threshMax)
{
    double* var = (double*)calloc(D, sizeof(double));
    const unsigned scale = 2;
    for (unsigned detInd = 0u; detInd < D; detInd++) {
        double varMax = data[detInd];
        double varMin = varMax;
        for (unsigned projInd = 0u; projInd < P - 1; projInd++) {
            var[detInd] += fabs((double)data[detInd + projInd * D] -
                                (double)data[detInd + (projInd + 1) * D]);
            if (data[detInd + projInd * D] > varMax) varMax = (double)data[detInd + projInd * D];
            if (data[detInd + projInd * D] < varMin) varMin = (double)data[detInd + projInd * D];
        }
        var[detInd] *= pow(varMax - varMin, (double)scale);
    }
    const int addNeighboursToFlickering = 2;
    for (int detectorSeg = 0; detectorSeg < 2; detectorSeg++) {
        for (unsigned i = 0u; i < D / 2; i++) {
            double thresh_segment = 0.0;
            for (int j = 0; j < 9; j++) {
                /* unsigned arithmetic: i - j wraps when j > i (:53-56) */
                int ind = (int)((i - (unsigned)j) % (D / 2));
                thresh_segment += ff[j] * var[(unsigned)ind + (unsigned)detectorSeg * (D / 2)];
                ind = (int)((i + (unsigned)j) % (D / 2));
                thresh_segment += ff[j] * var[(unsigned)ind + (unsigned)detectorSeg * (D / 2)];
            }
            unsigned detInd = (unsigned)detectorSeg * (D / 2) + i;
            if (var[detInd] < threshMin * thresh_segment) defect[detInd] = 1;
            if (var[detInd] > threshMax * thresh_segment) {
                for (int offset = -addNeighboursToFlickering; offset <= addNeighboursToFlickering; offset++)
                    defect[(detInd + (unsigned)offset) % D] = 1;
            }
        }
    }
    free(var);
}

/* interpolateDefectDetectors, interpolationFunctions.h:73-97. */
void orc_interp_defects(uint16_t* data, const int* defect, unsigned D, unsigned P)
{
    unsigned detID = 0;
    while (detID < D) {
        if (defect[detID]) {
            unsigned det0 = detID;
            while (detID < D && defect[detID] && defect[(detID + 1) % D]) detID++;
            for (unsigned i = det0; i <= detID; i++) {
                float w1 = (float)(((double)((float)i - (float)det0) + 1.0) /
                                   ((double)((float)detID - (float)det0) + 2.0));
                float w0 = (float)(1.0 - (double)w1);
                for (unsigned projId = 0u; projId < P; projId++) {
                    float v = w0 * (float)data[(det0 - 1) % D + projId * D] +
                              w1 * (float)data[(detID + 1) % D + projId * D];
                    data[(i % D) + projId * D] = (uint16_t)v;
                }
            }
        }
        detID++;
    }
}

/* Attenuation::init minus file IO (Attenuation.cu:177-217) with
 * computeAverage (:236-251) and computeDarkAverage (:219-234). */
void orc_reference_average(const orc_geom* g, uint16_t* ref, int R, float* avg_ref, float* avg_dark)
{
    const int D = g->numberOfFanDetectors, P = g->numberOfFanProjections;
    const int planes = g->numberOfPlanes;
    const size_t frame = (size_t)D * P;
    double ff[17];
    orc_defect_window(ff);
    int* flags = (int*)malloc(frame * sizeof(int));
    for (int i = 0; i < R * planes; i++) {                                      /* :191-196 */
        memset(flags, 0, frame * sizeof(int));
        orc_find_defects(ref + (size_t)i * frame, ff, flags, (unsigned)D, (unsigned)P,
                         g->threshMin, g->threshMax);
        orc_interp_defects(ref + (size_t)i * frame, flags, (unsigned)D, (unsigned)P);
    }
    free(flags);
    memset(avg_ref, 0, frame * planes * sizeof(float));
    const float factor = (float)(1.0 / (float)R);                               /* :240 */
    for (int i = 0; i < R; i++)
        for (int planeInd = 0; planeInd < planes; planeInd++)
            for (size_t index = 0; index < frame; index++)
                avg_ref[index + planeInd * frame] +=
                    (float)ref[(size_t)(i + planeInd) * frame + index] * factor;   /* :245-247 */
    /* :222-223: the dark factor is overwritten with 0, so the average is 0. */
    for (int k = 0; k < planes * D; k++) avg_dark[k] = 0.0f;
}

/* computeAttenuation, Attenuation.cu:416-444, temp = pow(10,-5) (:144). */
void orc_attenuation_f32(const orc_geom* g, const uint16_t* raw, int plane,
                         const float* avgReference, const float* avgDark,
                         const float* mask, float* out)
{
    const int D = g->numberOfFanDetectors, P = g->numberOfFanProjections;
    const float temp = (float)pow(10, -5);
    for (int y = 0; y < P; y++)
        for (int x = 0; x < D; x++) {
            const int sinoIndex = D * y + x;
            float numerator = (float)raw[sinoIndex] - avgDark[plane * D + x];
            float denominator = avgReference[(size_t)plane * D * P + sinoIndex] - avgDark[plane * D + x];
            if (numerator < temp) numerator = temp;
            if (denominator < temp) denominator = temp;
            out[sinoIndex] = -logf(numerator / denominator) * mask[sinoIndex];
        }
}

void orc_attenuation_f64(const orc_geom* g, const uint16_t* raw, int plane,
                         const float* avgReference, const float* avgDark,
                         const float* mask, double* out)
{
    const int D = g->numberOfFanDetectors, P = g->numberOfFanProjections;
    const double temp = pow(10, -5);
    for (int y = 0; y < P; y++)
        for (int x = 0; x < D; x++) {
            const int sinoIndex = D * y + x;
            double numerator = (double)raw[sinoIndex] - (double)avgDark[plane * D + x];
            double denominator = (double)avgReference[(size_t)plane * D * P + sinoIndex] -
                                 (double)avgDark[plane * D + x];
            if (numerator < temp) numerator = temp;
            if (denominator < temp) denominator = temp;
            out[sinoIndex] = -log(numerator / denominator) * (double)mask[sinoIndex];
        }
}

/* Reordering::createHashTable (Reordering.cu:180-192) and the reorder kernel
 * (:194-204) for `frames` consecutive frames: the table has P*D entries,
 * zero-initialised, and its first P*27*16 are filled in (projection, module,
 * detector) loop order; ordered[i] = raw[table[i]].  Returns -1 when D < 432,
 * where the reference's table indexes past the frame. */
int orc_reorder_u16(int P, int D, int frames, const uint16_t* raw, uint16_t* ordered)
{
    const int modules = 27, perModule = 16;
    if (D < modules * perModule) return -1;
    const size_t n = (size_t)P * D;
    int* table = (int*)calloc(n, sizeof(int));
    size_t i = 0;
    for (int projInd = 0; projInd < P; projInd++)
        for (int modInd = 0; modInd < modules; modInd++)
            for (int detInd = 0; detInd < perModule; detInd++)
                table[i++] = detInd + projInd * perModule + modInd * perModule * P;
    for (int f = 0; f < frames; f++)
        for (size_t e = 0; e < n; e++) ordered[f * n + e] = raw[f * n + (size_t)table[e]];
    free(table);
    return 0;
}
