/*
 * Oracle restatement of the Fan2Para stage (TEST INFRASTRUCTURE ONLY).
 *
 * Reference: RISA/risaLib/src/Fan2Para/Fan2Para.cu:251-480, 592-624 and
 *            RISA/risaLib/src/Fan2Para/cuda_kernels_fan2para.h:38-185.
 *
 * The host table code of the reference stores every value as float but mixes
 * in double literals, and calls unqualified sin/asin/cos/acos/sqrt/atan on
 * float arguments.  variant 0 evaluates those calls with the float overloads
 * (sinf, asinf, ...), variant 1 with the double overloads; every other
 * operation follows the C++ promotion rules of the original expression.
 */
#include "risa_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    int P2, D, Df, Pf, planes;
    float sourceOffset, rDetector, imageCenterX, imageCenterY, imageWidth;
    float rTarget[ORC_MAX_PLANES], deltaX[ORC_MAX_PLANES], deltaZ[ORC_MAX_PLANES];
    float sourceAngle[ORC_MAX_PLANES];
} f2p_params;

/* Fan2Para::readConfig, Fan2Para.cu:592-624 (float members). */
static void f2p_read(const orc_geom* g, f2p_params* p)
{
    p->P2 = g->numberOfParallelProjections * 2;                       /* :612 */
    p->D = g->numberOfParallelDetectors;
    p->Df = g->numberOfFanDetectors;
    p->Pf = g->numberOfFanProjections;
    p->planes = g->numberOfPlanes;
    p->sourceOffset = (float)g->sourceOffset;
    float detectorDiameter = (float)g->detectorDiameter;
    p->rDetector = (float)(detectorDiameter / 2.0);                  /* :611 */
    p->imageCenterX = (float)g->imageCenterX;
    p->imageCenterY = (float)g->imageCenterY;
    p->imageWidth = (float)g->imageWidth;
    for (int i = 0; i < p->planes && i < ORC_MAX_PLANES; i++) {
        float sourceDiam = (float)g->sourceDiameter[i];
        p->deltaX[i] = (float)g->deltaX[i];
        p->deltaZ[i] = (float)g->deltaZ[i];
        p->sourceAngle[i] = (float)g->sourceAngle[i];
        p->rTarget[i] = (float)(sourceDiam / 2.0);                  /* :618 */
    }
}

/* ellipse_kreis_uwe<float>, cuda_kernels_fan2para.h:38-62. */
float orc_ellipse_kreis_f32(float alpha, float DX, float DZ, float SourceRingDiam, int variant)
{
    float L, R, CA, eps, p1, p2, gamma, ae;
    if (variant == 0) L = sqrtf(DX * DX + DZ * DZ);
    else              L = (float)sqrt((double)(DX * DX + DZ * DZ));
    R = (float)(0.5 * (double)SourceRingDiam);
    if (variant == 0) CA = cosf(alpha);
    else              CA = (float)cos((double)alpha);

    eps = (float)((double)(L * L + R * DX * CA) /
                  ((double)L * sqrt((double)(L * L + R * R) + 2.0 * R * DX * CA)));
    if (variant == 0) eps = acosf(eps);
    else              eps = (float)acos((double)eps);

    p1 = (float)((double)(L * L - R * DX) / ((double)L * sqrt((double)(L * L + R * R) - 2.0 * R * DX)));
    p2 = (float)((double)(L * L + R * DX) / ((double)L * sqrt((double)(L * L + R * R) + 2.0 * R * DX)));
    if (variant == 0) gamma = (float)(0.5 * (double)(acosf(p1) - acosf(p2)));
    else              gamma = (float)(0.5 * (acos((double)p1) - acos((double)p2)));

    ae = (float)((double)(eps * CA + gamma) /
                 sqrt((double)(eps * eps) + 2.0 * eps * gamma * CA + (double)(gamma * gamma)));

    if ((double)alpha <= M_PI) {
        if (variant == 0) return acosf(ae);
        return (float)acos((double)ae);
    }
    if (variant == 0) return (float)(2.0 * M_PI - (double)acosf(ae));
    return (float)(2.0 * M_PI - acos((double)ae));
}

/* The same function evaluated entirely in double.  The two acos arguments
 * are cosines mathematically; rounding can push them past +-1 near alpha = 0
 * and pi, so they are clamped (the float form turns those into NaN, which
 * clears the ray flag). */
static double clamp1(double v) { return v > 1.0 ? 1.0 : (v < -1.0 ? -1.0 : v); }

double orc_ellipse_kreis_f64(double alpha, double DX, double DZ, double SourceRingDiam)
{
    double L = sqrt(DX * DX + DZ * DZ);
    double R = 0.5 * SourceRingDiam;
    double CA = cos(alpha);
    double eps = acos(clamp1((L * L + R * DX * CA) / (L * sqrt(L * L + R * R + 2.0 * R * DX * CA))));
    double p1 = (L * L - R * DX) / (L * sqrt(L * L + R * R - 2.0 * R * DX));
    double p2 = (L * L + R * DX) / (L * sqrt(L * L + R * R + 2.0 * R * DX));
    double gamma = 0.5 * (acos(p1) - acos(p2));
    double den = sqrt(eps * eps + 2.0 * eps * gamma * CA + gamma * gamma);
    double ae = den > 0.0 ? clamp1((eps * CA + gamma) / den) : 1.0;
    if (alpha <= M_PI) return acos(ae);
    return 2.0 * M_PI - acos(ae);
}

/* (s_i - L*sin(alpha_j - kappa)) / r, as float (variant 0) or as the double
 * quotient (variant 1).  Fan2Para.cu:344, :366, :389-390. */
static double ray_ratio(const f2p_params* p, int i, float alpha, float L, float kappa,
                        const float* s, float r, int variant)
{
    (void)p;
    if (variant == 0) {
        float v = (s[i] - L * sinf(alpha - kappa)) / r;
        return (double)v;
    }
    return ((double)s[i] - (double)L * sin((double)(alpha - kappa))) / (double)r;
}

static float f_asin(double x, int variant)
{
    if (variant == 0) return asinf((float)x);
    return (float)asin(x);
}

int orc_f2p_tables_build(const orc_geom* g, int variant, orc_f2p_tables* t)
{
    f2p_params p;
    f2p_read(g, &p);
    memset(t, 0, sizeof(*t));
    t->P2 = p.P2; t->D = p.D; t->Df = p.Df; t->Pf = p.Pf; t->planes = p.planes;
    const size_t n = (size_t)p.P2 * p.D * p.planes;
    t->theta = (float*)calloc(p.Pf, sizeof(float));
    t->gamma = (float*)calloc(p.Df, sizeof(float));
    t->s = (float*)calloc(p.D, sizeof(float));
    t->alpha = (float*)calloc(p.P2, sizeof(float));
    int32_t** iarr[] = {&t->thetaAfter1, &t->thetaAfter2, &t->thetaBefore1, &t->thetaBefore2,
                        &t->gammaAfter1, &t->gammaAfter2, &t->gammaBefore1, &t->gammaBefore2,
                        &t->ray1, &t->ray2};
    for (unsigned a = 0; a < sizeof(iarr) / sizeof(iarr[0]); a++) *iarr[a] = (int32_t*)calloc(n, sizeof(int32_t));
    float** farr[] = {&t->gammaGoal1, &t->gammaGoal2, &t->thetaGoal1, &t->thetaGoal2};
    for (unsigned a = 0; a < 4; a++) *farr[a] = (float*)calloc(n, sizeof(float));

    float* theta = t->theta; float* gam = t->gamma; float* s = t->s; float* alpha = t->alpha;

    /* :286-291 */
    for (int j = 0; j < p.Pf; j++) {
        theta[j] = (float)(j * (360.0 / p.Pf) - (double)p.sourceOffset);
        if (theta[j] < 0.0) theta[j] = (float)((double)theta[j] + 360.0);
        theta[j] = (float)(((2 * M_PI) / 360.0) * (double)theta[j]);
    }
    /* :293-297 */
    gam[0] = 0.0f;
    for (int j = 1; j < p.Df; j++) {
        gam[j] = (float)(j * (360.0 / (double)(float)p.Df));
        gam[j] = (float)(((2 * M_PI) / 360.0) * (double)gam[j]);
    }
    /* :299-301 */
    for (int j = 0; j < p.D; j++)
        s[j] = (float)(((-0.5) * (double)p.imageWidth) +
                       (((0.5 + j) * (double)p.imageWidth) / (double)(float)p.D));
    /* :303-309 -- runs j = P2 .. 1: alpha[0] keeps 0, the write to [P2] is
     * past the end of the reference vector and is not stored here. */
    for (int j = p.P2; j > 0; j--) {
        float a = (float)(j * (360.0 / (double)(float)p.P2));
        a = (float)(((2 * M_PI) / 360.0) * (double)a + M_PI / 2);
        if ((double)a > 2 * M_PI) a = (float)((double)a - (2.0 * M_PI));
        if (j < p.P2) alpha[j] = a;
    }

    /* :315-333 */
    float kappa = 0, L = 0, tb = 0;
    if (p.imageCenterY != 0) {
        float q = p.imageCenterY * p.imageCenterY + p.imageCenterX * p.imageCenterX;
        L = variant == 0 ? sqrtf(q) : (float)sqrt((double)q);
        if (p.imageCenterY < 0) tb = 1.0f;
        float r = p.imageCenterX / p.imageCenterY;
        kappa = variant == 0 ? (float)((double)atanf(r) + tb * M_PI) : (float)(atan((double)r) + tb * M_PI);
    } else if (p.imageCenterX != 0) {
        float q = p.imageCenterY * p.imageCenterY + p.imageCenterX * p.imageCenterX;
        L = variant == 0 ? sqrtf(q) : (float)sqrt((double)q);
        kappa = p.imageCenterX < 0 ? (float)(-M_PI / 2.) : (float)(M_PI / 2.);
    }

    const unsigned int parallelSize = (unsigned)p.D * (unsigned)p.P2;
    for (int k = 0; k < p.planes; k++) {
        const float rT = p.rTarget[k];
        /* :382-383 */
        const float lim1 = (float)(((360.0 - (double)p.sourceAngle[k]) / 2.0) / 180.0 * M_PI);
        const float lim2 = (float)((360.0 - ((360.0 - (double)p.sourceAngle[k]) / 2.0)) / 180.0 * M_PI);
        /* every (k, j, i) entry is independent (computeAngles keeps its
         * search state local), so the rows can be built in parallel */
#pragma omp parallel for schedule(dynamic, 8)
        for (int j = 0; j < p.P2; j++) {
            for (int i = 0; i < p.D; i++) {
                const size_t ind = (size_t)j * p.D + i + (size_t)k * parallelSize;    /* :343 */
                const float temp_range = (float)ray_ratio(&p, i, alpha[j], L, kappa, s, p.rDetector, variant);
                if (!(temp_range <= 1 && temp_range >= -1)) continue;                 /* :347 */

                /* computeAngles, :354-480 */
                float dif_best = (float)M_PI, dif = 0, best_x = 0, temp_1;
                temp_1 = f_asin(ray_ratio(&p, i, alpha[j], L, kappa, s, rT, variant), variant);
                float tg1 = alpha[j] - temp_1;
                if (tg1 < 0) tg1 = (float)((double)tg1 + 2.0 * M_PI);
                tg1 = orc_ellipse_kreis_f32(tg1, p.deltaX[k], p.deltaZ[k], 2 * rT, variant);
                float tg2 = (float)((double)(alpha[j] + temp_1) - M_PI);
                if (tg2 < 0) tg2 = (float)((double)tg2 + 2.0 * M_PI);
                tg2 = orc_ellipse_kreis_f32(tg2, p.deltaX[k], p.deltaZ[k], 2 * rT, variant);
                t->thetaGoal1[ind] = tg1;
                t->thetaGoal2[ind] = tg2;
                if (tg1 > lim1 && tg1 < lim2) t->ray1[ind] = 1;
                if (tg2 > lim1 && tg2 < lim2) t->ray2[ind] = 1;
                const float epsilon = f_asin(ray_ratio(&p, i, alpha[j], L, kappa, s, p.rDetector, variant), variant);

                if (t->ray1[ind]) {
                    float gg = (float)((double)(epsilon + alpha[j]) - 1.5 * M_PI);
                    if (gg < 0) gg = (float)((double)gg + 2.0 * M_PI);
                    if ((double)gg > 2 * M_PI) gg = (float)((double)gg - 2.0 * M_PI);
                    t->gammaGoal1[ind] = gg;
                    for (int x = 0; x < p.Pf; x++) {
                        if (tg1 <= theta[x]) {
                            dif = theta[x] - tg1;
                            if (dif < dif_best) { dif_best = dif; best_x = (float)x; }
                        }
                    }
                    if (best_x == 0) { t->thetaBefore1[ind] = p.Pf - 1; t->thetaAfter1[ind] = (int)best_x; }
                    else { t->thetaBefore1[ind] = (int)(best_x - 1); t->thetaAfter1[ind] = (int)best_x; }
                    for (int x = 0; x < p.Df; x++) {
                        if (gg <= gam[x]) {
                            t->gammaBefore1[ind] = (x == 0) ? p.Df - 1 : x - 1;
                            t->gammaAfter1[ind] = x;
                            break;
                        }
                    }
                    if (gg > gam[p.Df - 1]) { t->gammaBefore1[ind] = p.Df - 1; t->gammaAfter1[ind] = 0; }
                }
                if (t->ray2[ind]) {
                    dif_best = (float)M_PI;    /* best_x is NOT reset (:438) */
                    float gg = (float)((double)(-epsilon + alpha[j]) - (M_PI / 2.0));
                    if (gg < 0) gg = (float)((double)gg + 2.0 * M_PI);
                    t->gammaGoal2[ind] = gg;
                    for (int x = 0; x < p.Pf; x++) {
                        if (tg2 <= theta[x]) {
                            dif = theta[x] - tg2;
                            if (dif < dif_best) { dif_best = dif; best_x = (float)x; }
                        }
                    }
                    if (best_x == 0) { t->thetaBefore2[ind] = p.Pf - 1; t->thetaAfter2[ind] = (int)best_x; }
                    else { t->thetaBefore2[ind] = (int)(best_x - 1); t->thetaAfter2[ind] = (int)best_x; }
                    for (int x = 0; x < p.Df; x++) {
                        if (gg <= gam[x]) {
                            t->gammaBefore2[ind] = (x == 0) ? p.Df - 1 : x - 1;
                            t->gammaAfter2[ind] = x;
                            break;
                        }
                    }
                    if (gg > gam[p.Df - 1]) { t->gammaBefore2[ind] = p.Df - 1; t->gammaAfter2[ind] = 0; }
                }
            }
        }
    }
    return 0;
}

void orc_f2p_tables_free(orc_f2p_tables* t)
{
    void* ptrs[] = {t->theta, t->gamma, t->s, t->alpha, t->thetaAfter1, t->thetaAfter2,
                    t->thetaBefore1, t->thetaBefore2, t->gammaAfter1, t->gammaAfter2,
                    t->gammaBefore1, t->gammaBefore2, t->gammaGoal1, t->gammaGoal2,
                    t->thetaGoal1, t->thetaGoal2, t->ray1, t->ray2};
    for (unsigned a = 0; a < sizeof(ptrs) / sizeof(ptrs[0]); a++) free(ptrs[a]);
    memset(t, 0, sizeof(*t));
}

/* interpolation kernel, cuda_kernels_fan2para.h:64-177 (the zero fill of
 * setValue, :179-185, is the memset).  The two atomicAdd's per output of the
 * 360->180 fold commute exactly (0 + a + b). */
void orc_f2p_gather_f32(const orc_geom* g, const orc_f2p_tables* t, int k,
                        const float* fan, float* par)
{
    f2p_params p;
    f2p_read(g, &p);
    const int D = p.D, P2 = p.P2, Df = p.Df;
    memset(par, 0, sizeof(float) * (size_t)D * (P2 / 2));
    const float* Teta = t->theta; const float* Gamma = t->gamma;
    for (int j = 0; j < P2; j++) {
        for (int i = 0; i < D; i++) {
            float WZiel_1 = 0, WZiel_2 = 0, WZiel_end = 0;
            const size_t ind = (size_t)j * D + i + (size_t)k * P2 * D;
            const float temp_1 = t->s[i] / p.rDetector;
            if (temp_1 <= 1 && temp_1 >= -1) {
                if (t->ray1[ind]) {
                    const int tb = t->thetaBefore1[ind], ta = t->thetaAfter1[ind];
                    const int gb = t->gammaBefore1[ind], ga = t->gammaAfter1[ind];
                    float W1 = fan[tb * Df + gb], W2 = fan[tb * Df + ga];
                    float W3 = fan[ta * Df + gb], W4 = fan[ta * Df + ga];
                    float V1 = W1 + ((t->thetaGoal1[ind] - Teta[tb]) / (Teta[ta] - Teta[tb])) * (W3 - W1);
                    float V2 = W2 + ((t->thetaGoal1[ind] - Teta[tb]) / (Teta[ta] - Teta[tb])) * (W4 - W2);
                    WZiel_1 = V1 + ((t->gammaGoal1[ind] - Gamma[gb]) / (Gamma[ga] - Gamma[gb])) * (V2 - V1);
                }
                if (t->ray2[ind]) {
                    const int tb = t->thetaBefore2[ind], ta = t->thetaAfter2[ind];
                    const int gb = t->gammaBefore2[ind], ga = t->gammaAfter2[ind];
                    float W1 = fan[tb * Df + gb], W2 = fan[tb * Df + ga];
                    float W3 = fan[ta * Df + gb], W4 = fan[ta * Df + ga];
                    float V1 = W1 + ((t->thetaGoal2[ind] - Teta[tb]) / (Teta[ta] - Teta[tb])) * (W3 - W1);
                    float V2 = W2 + ((t->thetaGoal2[ind] - Teta[tb]) / (Teta[ta] - Teta[tb])) * (W4 - W2);
                    WZiel_2 = V1 + ((t->gammaGoal2[ind] - Gamma[gb]) / (Gamma[ga] - Gamma[gb])) * (V2 - V1);
                }
                const int r1 = t->ray1[ind], r2 = t->ray2[ind];
                if (r1 + r2 > 0)
                    WZiel_end = (float)r1 / ((float)r1 + (float)r2) * WZiel_1 +
                                (float)r2 / ((float)r1 + (float)r2) * WZiel_2;
            }
            const int address = j * D;
            const int parallelSize = P2 * D;
            const int detectorSize_2 = D / 2;
            int mirrorOffset;
            if (i < detectorSize_2) mirrorOffset = D - i - 1;
            else mirrorOffset = -(i % detectorSize_2) + detectorSize_2 - 1;
            const float add = (float)((double)WZiel_end * 0.5);
            if (j < P2 / 2) par[address + i] += add;
            else par[address - parallelSize / 2 + mirrorOffset] += add;
        }
    }
}

/* 2 pi representative of a nearest to `near` (the float goal angle), so the
 * double form keeps the float form's wrap decisions. */
static double near_rep(double a, double near)
{
    return a + 2.0 * M_PI * nearbyint((near - a) / (2.0 * M_PI));
}

/* FP64 form: coordinates, goal angles and weights in double, integer
 * decisions (indices, ray flags, wraps) from the faithful table t. */
void orc_f2p_gather_f64(const orc_geom* g, const orc_f2p_tables* t, int k,
                        const double* fan, double* par)
{
    const int D = t->D, P2 = t->P2, Df = t->Df, Pf = t->Pf;
    double L = 0.0, kappa = 0.0;
    if (g->imageCenterY != 0) {
        L = sqrt(g->imageCenterY * g->imageCenterY + g->imageCenterX * g->imageCenterX);
        kappa = atan(g->imageCenterX / g->imageCenterY) + (g->imageCenterY < 0 ? M_PI : 0.0);
    } else if (g->imageCenterX != 0) {
        L = fabs(g->imageCenterX);
        kappa = g->imageCenterX < 0 ? -M_PI / 2 : M_PI / 2;
    }
    double* theta = (double*)malloc(sizeof(double) * Pf);
    double* gam = (double*)malloc(sizeof(double) * Df);
    for (int j = 0; j < Pf; j++) {
        double v = j * (360.0 / Pf) - g->sourceOffset;
        if (v < 0.0) v += 360.0;
        theta[j] = (2 * M_PI / 360.0) * v;
    }
    gam[0] = 0.0;
    for (int j = 1; j < Df; j++) gam[j] = (2 * M_PI / 360.0) * (j * (360.0 / Df));
    const double rD = g->detectorDiameter / 2.0;
    const double rT = g->sourceDiameter[k] / 2.0;
    memset(par, 0, sizeof(double) * (size_t)D * (P2 / 2));
    for (int j = 0; j < P2; j++) {
        double alpha = 0.0;                            /* alpha[0] = 0 as in the table */
        if (j > 0) {
            alpha = (2 * M_PI / 360.0) * (j * (360.0 / P2)) + M_PI / 2;
            if (alpha > 2 * M_PI) alpha -= 2.0 * M_PI;
        }
        for (int i = 0; i < D; i++) {
            const size_t ind = (size_t)j * D + i + (size_t)k * P2 * D;
            const double s = -0.5 * g->imageWidth + ((0.5 + i) * g->imageWidth) / D;
            double Z1 = 0, Z2 = 0, Zend = 0;
            const int r1 = t->ray1[ind], r2 = t->ray2[ind];
            if (r1 + r2 > 0) {
                const double off = s - L * sin(alpha - kappa);
                const double t1 = asin(off / rT);
                const double eps = asin(off / rD);
                if (r1) {
                    double tg = alpha - t1;
                    if (tg < 0) tg += 2.0 * M_PI;
                    tg = orc_ellipse_kreis_f64(tg, g->deltaX[k], g->deltaZ[k], 2 * rT);
                    double gg = eps + alpha - 1.5 * M_PI;
                    if (gg < 0) gg += 2.0 * M_PI;
                    if (gg > 2 * M_PI) gg -= 2.0 * M_PI;
                    tg = near_rep(tg, t->thetaGoal1[ind]);
                    gg = near_rep(gg, t->gammaGoal1[ind]);
                    const int tb = t->thetaBefore1[ind], ta = t->thetaAfter1[ind];
                    const int gb = t->gammaBefore1[ind], ga = t->gammaAfter1[ind];
                    const double ft = (tg - theta[tb]) / (theta[ta] - theta[tb]);
                    const double fg = (gg - gam[gb]) / (gam[ga] - gam[gb]);
                    const double V1 = fan[tb * Df + gb] + ft * (fan[ta * Df + gb] - fan[tb * Df + gb]);
                    const double V2 = fan[tb * Df + ga] + ft * (fan[ta * Df + ga] - fan[tb * Df + ga]);
                    Z1 = V1 + fg * (V2 - V1);
                }
                if (r2) {
                    double tg = alpha + t1 - M_PI;
                    if (tg < 0) tg += 2.0 * M_PI;
                    tg = orc_ellipse_kreis_f64(tg, g->deltaX[k], g->deltaZ[k], 2 * rT);
                    double gg = -eps + alpha - M_PI / 2.0;
                    if (gg < 0) gg += 2.0 * M_PI;
                    tg = near_rep(tg, t->thetaGoal2[ind]);
                    gg = near_rep(gg, t->gammaGoal2[ind]);
                    const int tb = t->thetaBefore2[ind], ta = t->thetaAfter2[ind];
                    const int gb = t->gammaBefore2[ind], ga = t->gammaAfter2[ind];
                    const double ft = (tg - theta[tb]) / (theta[ta] - theta[tb]);
                    const double fg = (gg - gam[gb]) / (gam[ga] - gam[gb]);
                    const double V1 = fan[tb * Df + gb] + ft * (fan[ta * Df + gb] - fan[tb * Df + gb]);
                    const double V2 = fan[tb * Df + ga] + ft * (fan[ta * Df + ga] - fan[tb * Df + ga]);
                    Z2 = V1 + fg * (V2 - V1);
                }
                Zend = ((double)r1 * Z1 + (double)r2 * Z2) / (double)(r1 + r2);
            }
            const int detectorSize_2 = D / 2;
            int mirrorOffset;
            if (i < detectorSize_2) mirrorOffset = D - i - 1;
            else mirrorOffset = -(i % detectorSize_2) + detectorSize_2 - 1;
            if (j < P2 / 2) par[j * D + i] += 0.5 * Zend;
            else par[j * D - (P2 * D) / 2 + mirrorOffset] += 0.5 * Zend;
        }
    }
    free(theta);
    free(gam);
}
