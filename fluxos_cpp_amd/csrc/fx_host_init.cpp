// fx_host_init.cpp -- host input layer of libfluxos_hip.so: the bed and weir
// mask from a parsed DEM (read_functions.cpp:145-225) and the cold-start
// initial condition with its ghost-cell treatment (initiate.cpp:27-231).
// Plain C++; no device code.  The arithmetic keeps the reference's dtypes
// (the DEM is read as float, averaging of weir neighbours is done in float,
// the vertex-to-centre average in double) because the initial bed decides
// every later step.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fluxos_hip.h"

int fx_fail(int code, const std::string &msg);

extern "C" int fx_read_geo(int32_t nrows, int32_t ncols, double nodata, const float *dem,
                           double ks_input, double *zb, double *ks, float *weir)
{
    if (nrows < 3 || ncols < 3 || !dem || !zb || !ks || !weir)
        return fx_fail(FX_ERR_ARG, "fx_read_geo: bad arguments");
    const int64_t mr = nrows + 2, mc = ncols + 2;
    auto at = [mr](int r, int c) { return (size_t)r + (size_t)c * (size_t)mr; };
    // zb_raw: zero outside the parsed values; first file line is irow = NROWS
    std::vector<float> raw((size_t)(mr * mc), 0.0f);
    for (int line = 0; line < nrows; line++) {
        const int irow = nrows - line;
        for (int j = 0; j < ncols; j++) raw[at(irow, j + 1)] = dem[(size_t)line * ncols + j];
    }
    auto valid = [nodata](float v) { return (double)v != nodata && v != 0.0f; };
    for (int icol = 1; icol <= ncols; icol++) {
        for (int irow = 1; irow <= nrows; irow++) {
            const float zbp = raw[at(irow, icol)];
            float corr;
            if ((double)zbp == nodata || zbp == 0.0f) {
                // weir cell: bed = mean of the valid 4-neighbours (float sums)
                int n = 0;
                corr = 0.0f;
                weir[at(irow, icol)] = 1.0f;
                const float zn = raw[at(irow + 1, icol)], zs = raw[at(irow - 1, icol)];
                const float ze = raw[at(irow, icol + 1)], zw = raw[at(irow, icol - 1)];
                if (valid(zn) && irow < nrows) { corr += zn; n++; }
                if (valid(zs) && irow > 1) { corr += zs; n++; }
                if (valid(ze) && icol < ncols) { corr += ze; n++; }
                if (valid(zw) && icol > 1) { corr += zw; n++; }
                if (corr == 0.0f) corr = zbp;
                else corr /= n;
            } else {
                weir[at(irow, icol)] = 0.0f;
                corr = std::fabs(zbp);
            }
            zb[at(irow, icol)] = (double)corr;
            ks[at(irow, icol)] = ks_input;
        }
    }
    return FX_OK;
}

extern "C" int fx_initiation(int32_t nrows, int32_t ncols, int32_t ade_solver, double *zb,
                             double *z, double *h, float *ldry, double *conc0)
{
    if (nrows < 3 || ncols < 3 || !zb || !z || !h || !ldry)
        return fx_fail(FX_ERR_ARG, "fx_initiation: bad arguments");
    const int64_t mr = nrows + 2;
    auto at = [mr](int r, int c) { return (size_t)r + (size_t)c * (size_t)mr; };
    const int nr1 = nrows + 1, nc1 = ncols + 1;
    // copies of the first/last row and column before averaging (initiate.cpp:43-55);
    // the saved first column is taken before rows 0 and NROWS+1 are filled
    std::vector<double> col1((size_t)mr, 0.0);
    for (int irow = 0; irow <= nr1; irow++) {
        col1[irow] = zb[at(irow, 1)];
        zb[at(irow, 0)] = zb[at(irow, 1)];
        zb[at(irow, nc1)] = zb[at(irow, ncols)];
    }
    for (int icol = 0; icol <= nc1; icol++) {
        zb[at(0, icol)] = zb[at(1, icol)];
        zb[at(nr1, icol)] = zb[at(nrows, icol)];
    }
    zb[at(0, nc1)] = zb[at(1, ncols)];
    zb[at(nr1, nc1)] = zb[at(nrows, ncols)];
    // cell value = mean of the four surrounding vertex values that are not
    // +-99999 (initiate.cpp:57-90); col1 rolls along as the previous column
    for (int icol = 1; icol <= ncols; icol++) {
        double sw = col1[1];
        double nw = zb[at(1, icol + 1)];
        col1[1] = nw;
        for (int irow = 1; irow <= nrows; irow++) {
            const double se = col1[irow + 1];
            const double ne = zb[at(irow + 1, icol + 1)];
            double sum = 0;
            unsigned cnt = 0;
            const double v[4] = {sw, se, nw, ne};
            for (double x : v)
                if (std::fabs(x) != 99999) { sum = sum + x; cnt = cnt + 1; }
            zb[at(irow, icol)] = (sum) / cnt;
            col1[irow + 1] = ne;
            sw = se;
            nw = ne;
        }
    }
    // cold start: dry, still water at the bed (initiate.cpp:166-187)
    for (int icol = 1; icol <= ncols; icol++) {
        for (int irow = 1; irow <= nrows; irow++) {
            const size_t p = at(irow, icol);
            h[p] = 0.0;
            z[p] = zb[p];
            ldry[p] = 1.0f;
            if (conc0) conc0[p] = ade_solver ? 0.0 : -1.0;
        }
    }
    // ghost rows then ghost columns by linear extrapolation (initiate.cpp:192-228)
    for (int icol = 0; icol <= nc1; icol++) {
        zb[at(0, icol)] = 1.5 * zb[at(1, icol)] - .5 * zb[at(2, icol)];
        zb[at(nr1, icol)] = 1.5 * zb[at(nrows, icol)] - .5 * zb[at(nrows - 1, icol)];
        z[at(0, icol)] = 1.5 * z[at(1, icol)] - .5 * z[at(2, icol)];
        z[at(nr1, icol)] = 1.5 * z[at(nrows, icol)] - .5 * z[at(nrows - 1, icol)];
        h[at(0, icol)] = std::fmax(0.0, z[at(0, icol)] - zb[at(0, icol)]);
        h[at(nr1, icol)] = std::fmax(0.0, z[at(nr1, icol)] - zb[at(nr1, icol)]);
    }
    for (int irow = 0; irow <= nr1; irow++) {
        zb[at(irow, 0)] = 1.5 * zb[at(irow, 1)] - .5 * zb[at(irow, 2)];
        zb[at(irow, nc1)] = 1.5 * zb[at(irow, ncols)] - .5 * zb[at(irow, ncols - 1)];
        z[at(irow, 0)] = 1.5 * z[at(irow, 1)] - .5 * z[at(irow, 2)];
        z[at(irow, nc1)] = 1.5 * z[at(irow, ncols)] - .5 * z[at(irow, ncols - 1)];
        h[at(irow, 0)] = std::fmax(0.0, z[at(irow, 0)] - zb[at(irow, 0)]);
        h[at(irow, nc1)] = std::fmax(0.0, z[at(irow, nc1)] - zb[at(irow, nc1)]);
    }
    return FX_OK;
}
