// fx_host_init.cpp -- host input layer of libfluxos_hip.so: the bed and weir
// mask from a parsed DEM (read_functions.cpp:145-225) and the initial
// condition, cold or from a Results file, with its ghost-cell treatment
// (initiate.cpp:27-231).
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

namespace {

struct grid_ix {
    int64_t mr;
    size_t operator()(int r, int c) const { return (size_t)r + (size_t)c * (size_t)mr; }
};

// ghost copies of the raw bed, then cell value = mean of the four
// surrounding vertex values that are not +-99999 (initiate.cpp:43-90); the
// saved first column is taken before rows 0 and NROWS+1 are filled, and
// rolls along as the previous column
void bed_to_centres(int32_t nrows, int32_t ncols, double *zb)
{
    const grid_ix at{(int64_t)nrows + 2};
    const int nr1 = nrows + 1, nc1 = ncols + 1;
    std::vector<double> col1((size_t)at.mr, 0.0);
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
}

// ghost rows then ghost columns by linear extrapolation (initiate.cpp:192-228);
// qx, qy may be NULL (their ghost values are set to 0)
void ghost_extrapolation(int32_t nrows, int32_t ncols, double *zb, double *z, double *h,
                         double *qx, double *qy)
{
    const grid_ix at{(int64_t)nrows + 2};
    const int nr1 = nrows + 1, nc1 = ncols + 1;
    for (int icol = 0; icol <= nc1; icol++) {
        zb[at(0, icol)] = 1.5 * zb[at(1, icol)] - .5 * zb[at(2, icol)];
        zb[at(nr1, icol)] = 1.5 * zb[at(nrows, icol)] - .5 * zb[at(nrows - 1, icol)];
        z[at(0, icol)] = 1.5 * z[at(1, icol)] - .5 * z[at(2, icol)];
        z[at(nr1, icol)] = 1.5 * z[at(nrows, icol)] - .5 * z[at(nrows - 1, icol)];
        h[at(0, icol)] = std::fmax(0.0, z[at(0, icol)] - zb[at(0, icol)]);
        h[at(nr1, icol)] = std::fmax(0.0, z[at(nr1, icol)] - zb[at(nr1, icol)]);
        if (qx) qx[at(0, icol)] = 0.0;
        if (qy) qy[at(nr1, icol)] = 0.0;
    }
    for (int irow = 0; irow <= nr1; irow++) {
        zb[at(irow, 0)] = 1.5 * zb[at(irow, 1)] - .5 * zb[at(irow, 2)];
        zb[at(irow, nc1)] = 1.5 * zb[at(irow, ncols)] - .5 * zb[at(irow, ncols - 1)];
        z[at(irow, 0)] = 1.5 * z[at(irow, 1)] - .5 * z[at(irow, 2)];
        z[at(irow, nc1)] = 1.5 * z[at(irow, ncols)] - .5 * z[at(irow, ncols - 1)];
        h[at(irow, 0)] = std::fmax(0.0, z[at(irow, 0)] - zb[at(irow, 0)]);
        h[at(irow, nc1)] = std::fmax(0.0, z[at(irow, nc1)] - zb[at(irow, nc1)]);
        if (qx) qx[at(irow, 0)] = 0.0;
        if (qy) qy[at(irow, nc1)] = 0.0;
    }
}

} // namespace

extern "C" int fx_initiation(int32_t nrows, int32_t ncols, int32_t ade_solver, double *zb,
                             double *z, double *h, float *ldry, double *conc0)
{
    if (nrows < 3 || ncols < 3 || !zb || !z || !h || !ldry)
        return fx_fail(FX_ERR_ARG, "fx_initiation: bad arguments");
    const grid_ix at{(int64_t)nrows + 2};
    bed_to_centres(nrows, ncols, zb);
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
    ghost_extrapolation(nrows, ncols, zb, z, h, nullptr, nullptr);
    return FX_OK;
}

extern "C" int fx_initiation_restart(int32_t nrows, int32_t ncols, int32_t ade_solver,
                                     double soil_conc_bckgrd, const double *table, int64_t table_rows,
                                     fx_init_grids *g)
{
    if (nrows < 3 || ncols < 3 || !g || !g->zb || !g->z || !g->h || !g->qx || !g->qy || !g->us ||
        !g->ux || !g->uy || !g->conc0 || !g->soil_mass || !g->twet || !g->ldry || table_rows < 0 ||
        (table_rows > 0 && !table))
        return fx_fail(FX_ERR_ARG, "fx_initiation_restart: bad arguments");
    const grid_ix at{(int64_t)nrows + 2};
    const int64_t mc = (int64_t)ncols + 2;
    const size_t mm = (size_t)at.mr * (size_t)mc;
    // rows naming a cell outside the grid are refused before anything changes
    for (int64_t a = 0; a < table_rows; a++) {
        const unsigned r = (unsigned)(int)table[a * 16], c = (unsigned)(int)table[a * 16 + 1];
        if (r >= (unsigned)at.mr || c >= (unsigned)mc)
            return fx_fail(FX_ERR_ARG, "Results row " + std::to_string(a + 1) + " names a cell outside the grid");
    }
    // the reference's freshly constructed matrices
    for (double *v : {g->z, g->h, g->qx, g->qy, g->us, g->ux, g->uy, g->conc0, g->soil_mass, g->twet})
        std::fill(v, v + mm, 0.0);
    std::fill(g->ldry, g->ldry + mm, 0.0f);
    bed_to_centres(nrows, ncols, g->zb);
    // depth from the (zero) surface, discharge from the (zero) velocity and
    // the background soil mass, where |zb| != 99999 (initiate.cpp:93-107)
    for (int icol = 1; icol <= ncols; icol++) {
        for (int irow = 1; irow <= nrows; irow++) {
            const size_t p = at(irow, icol);
            if (std::fabs(g->zb[p]) != 99999) {
                g->h[p] = std::fmax(g->z[p] - g->zb[p], 0.0);
                g->z[p] = g->zb[p] + g->h[p];
                g->qx[p] = g->ux[p] * g->h[p];
                g->qy[p] = g->uy[p] * g->h[p];
                g->soil_mass[p] = soil_conc_bckgrd;
            }
        }
    }
    // the file's rows (initiate.cpp:132-153)
    for (int64_t a = 0; a < table_rows; a++) {
        const double *r = table + a * 16;
        const size_t p = at((int)r[0], (int)r[1]);
        g->z[p] = r[4];
        g->h[p] = r[5];
        g->ux[p] = r[6];
        g->uy[p] = r[7];
        g->qx[p] = r[8];
        g->qy[p] = r[9];
        g->us[p] = r[10];
        g->conc0[p] = ade_solver ? std::fmax(r[11], 0.0) : -1.0;
        g->soil_mass[p] = r[12];
        g->twet[p] = r[15];
        g->ldry[p] = 0.0f;
    }
    ghost_extrapolation(nrows, ncols, g->zb, g->z, g->h, g->qx, g->qy);
    return FX_OK;
}
