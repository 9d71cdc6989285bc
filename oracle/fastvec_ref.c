/*
 * C restatement of the reference's fastVEC kernel -- TEST INFRASTRUCTURE ONLY
 * (the oracle's exact-order form; also the CPU baseline for the vector field).
 *
 * Follows VESPER_1_prob_new.py:180-257 (doFastVEC) and :121-142 (calc) in the
 * reference's own loop order, summation order and float64 arithmetic, with
 * no FMA contraction (built with -ffp-contract=off, as Numba's default
 * non-fastmath LLVM IR does not contract either).
 *
 * Layouts: src[sx][sy][sz] float32 (the reference's [x][y][z] data),
 * out_data[n][n][n] float64, out_vec[n][n][n][3] float32.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

/* Output planes x in [x0, x1) only (each voxel is independent of the
 * others); the rest of out_data / out_vec is left zero. */
int fastvec_ref_planes(const float *src, int sx, int sy, int sz,
                       double src_width, const double *src_orig,
                       int n, double dest_width, const double *dest_orig,
                       double dreso, double *out_data, float *out_vec, int x0, int x1)
{
    const double gstep = src_width;                 /* :183 */
    double fs = (dreso / gstep) * 0.5;              /* :184 */
    fs = fs * fs;                                   /* :185 */
    const double fsiv = 1.0 / fs;                   /* :186 */
    const double fmaxd = (dreso / gstep) * 2.0;     /* :187 */
    const int dims[3] = {sx, sy, sz};
    const size_t n3 = (size_t)n * n * n;
    memset(out_data, 0, n3 * sizeof(double));
    memset(out_vec, 0, n3 * 3 * sizeof(float));

    if (x0 < 0) x0 = 0;
    if (x1 > n) x1 = n;
    for (int x = x0; x < x1; ++x)
    for (int y = 0; y < n; ++y)
    for (int z = 0; z < n; ++z) {
        const int xyz[3] = {x, y, z};
        double pos[3];
        int skip = 0;
        for (int a = 0; a < 3; ++a) {                /* :196 */
            pos[a] = (((double)xyz[a] * dest_width + dest_orig[a]) - src_orig[a]) / gstep;
            if (pos[a] < 0 || pos[a] >= (double)dims[a]) skip = 1;     /* :200-208 */
        }
        if (skip) continue;
        if (src[((size_t)(int)pos[0] * sy + (int)pos[1]) * sz + (int)pos[2]] == 0.0f)
            continue;                                                 /* :210 */
        int stp[3], endp[3];
        for (int a = 0; a < 3; ++a) {
            stp[a] = (int32_t)(pos[a] - fmaxd);                       /* :214 */
            if (stp[a] < 0) stp[a] = 0;
            endp[a] = (int32_t)(pos[a] + fmaxd + 1);                  /* :225 */
            if (endp[a] >= dims[a]) endp[a] = dims[a];
        }
        double dtotal = 0.0, p2[3] = {0.0, 0.0, 0.0};                /* calc :121-142 */
        for (int xp = stp[0]; xp < endp[0]; ++xp) {
            double rx = (double)xp - pos[0];
            rx = rx * rx;
            for (int yp = stp[1]; yp < endp[1]; ++yp) {
                double ry = (double)yp - pos[1];
                ry = ry * ry;
                const float *row = src + ((size_t)xp * sy + yp) * sz;
                for (int zp = stp[2]; zp < endp[2]; ++zp) {
                    double rz = (double)zp - pos[2];
                    rz = rz * rz;
                    const double d2 = rx + ry + rz;
                    const double v = (double)row[zp] * exp(-1.5 * d2 * fsiv);
                    dtotal += v;
                    p2[0] += v * xp;
                    p2[1] += v * yp;
                    p2[2] += v * zp;
                }
            }
        }
        const size_t o = ((size_t)x * n + y) * n + z;
        out_data[o] = dtotal;                                         /* :237 */
        if (dtotal == 0) continue;
        const double rd = 1.0 / dtotal;
        double t[3];
        for (int a = 0; a < 3; ++a) t[a] = p2[a] * rd - pos[a];        /* :244-246 */
        double dvec = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        if (dvec == 0) dvec = 1.0;
        const double rdvec = 1.0 / dvec;
        for (int a = 0; a < 3; ++a) out_vec[o * 3 + a] = (float)(t[a] * rdvec);   /* :255 */
    }
    return 0;
}

int fastvec_ref(const float *src, int sx, int sy, int sz,
                double src_width, const double *src_orig,
                int n, double dest_width, const double *dest_orig,
                double dreso, double *out_data, float *out_vec)
{
    return fastvec_ref_planes(src, sx, sy, sz, src_width, src_orig, n, dest_width, dest_orig, dreso,
                              out_data, out_vec, 0, n);
}
