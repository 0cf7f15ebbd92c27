// fx_io.cpp -- the reference's file formats at the boundary of the step:
// the ESRI-ASCII DEM (read_functions.cpp:112-164, common.cpp:110-161), the
// forcing CSV files (read_functions.cpp:237-326), the Results/<t>.txt
// writer (write_results.cpp:30-102) and its reader for a restart
// (initiate.cpp:109-156, common.cpp:69-106).  Host code only; no GPU needed.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dirent.h>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/fluxos_hip.h"

int fx_fail(int code, const std::string &msg);   // fx_capi.hip

namespace {

// first whitespace-separated token of a header line that parses as an int
// (getIntNumberFromString, common.cpp:38-51: "2.5" reads as 2)
bool first_int(const std::string &line, long long *v)
{
    std::istringstream ss(line);
    std::string tok;
    while (ss >> tok) {
        std::istringstream t(tok);
        long long x;
        if (t >> x) { *v = x; return true; }
    }
    return false;
}

bool first_double(const std::string &line, double *v)
{
    std::istringstream ss(line);
    std::string tok;
    while (ss >> tok) {
        std::istringstream t(tok);
        double x;
        if (t >> x) { *v = x; return true; }
    }
    return false;
}

std::string upper5(const std::string &s)
{
    std::string u = s.substr(0, 5);
    for (auto &c : u) c = (char)toupper((unsigned char)c);
    return u;
}

} // namespace

extern "C" {

int fx_read_dem(const char *path, fx_dem_header *hdr, float *dem)
{
    if (!path || !hdr) return fx_fail(FX_ERR_ARG, "path or header is NULL");
    std::ifstream f(path);
    if (!f.is_open()) return fx_fail(FX_ERR_IO, std::string("cannot open DEM file ") + path);
    std::string line[6];
    for (int i = 0; i < 6; i++)
        if (!std::getline(f, line[i])) return fx_fail(FX_ERR_IO, std::string("DEM header is short: ") + path);
    // NROWS / NCOLS from the first two lines, in either order (common.cpp:134-149)
    long long nr = -1, nc = -1, v;
    for (int i = 0; i < 2; i++) {
        if (!first_int(line[i], &v)) return fx_fail(FX_ERR_IO, "DEM header: no number on line " + std::to_string(i + 1));
        const std::string key = upper5(line[i]);
        if (key == "NROWS") nr = v;
        else if (key == "NCOLS") nc = v;
    }
    if (nr < 3 || nc < 3) return fx_fail(FX_ERR_IO, "DEM header: NROWS/NCOLS missing or < 3");
    double xll = 0.0, yll = 0.0;
    long long dxy = 0, nodata = 0;
    if (!first_double(line[2], &xll) || !first_double(line[3], &yll) || !first_int(line[4], &dxy) ||
        !first_int(line[5], &nodata))
        return fx_fail(FX_ERR_IO, "DEM header: XLLCORNER/YLLCORNER/CELLSIZE/NODATA_VALUE unreadable");
    if (dxy <= 0) return fx_fail(FX_ERR_IO, "DEM header: CELLSIZE must be a positive integer");
    hdr->nrows = (int32_t)nr;
    hdr->ncols = (int32_t)nc;
    hdr->xllcorner = xll;
    hdr->yllcorner = yll;
    hdr->cellsize = (int32_t)dxy;
    hdr->nodata = (double)nodata;
    if (!dem) return FX_OK;
    // data lines in file order (the first line is irow = NROWS,
    // read_functions.cpp:145-163); tokens that are not numbers are skipped
    // and leave their value at 0, like the reference's untouched zb_raw
    std::fill(dem, dem + (size_t)nr * nc, 0.0f);
    std::string s;
    long long r = 0;
    while (r < nr && std::getline(f, s)) {
        std::istringstream ss(s);
        std::string tok;
        long long c = 0;
        while (ss >> tok) {
            std::istringstream t(tok);
            float x;
            if (c < nc && (t >> x)) dem[(size_t)r * nc + c] = x;
            c++;
        }
        r++;
    }
    return FX_OK;
}

int fx_read_forcing(const char *path, int32_t max_rows, int32_t max_cols, float *table, int32_t *nrows,
                    int32_t *ncols)
{
    if (!path || !nrows || !ncols) return fx_fail(FX_ERR_ARG, "path or outputs are NULL");
    std::ifstream f(path);
    if (!f.is_open()) return fx_fail(FX_ERR_IO, std::string("cannot open forcing file ") + path);
    std::string s;
    std::vector<std::vector<double>> rows;
    bool first = true;
    while (std::getline(f, s)) {
        if (first) { first = false; continue; }         // header line (read_functions.cpp:258,306)
        if (s.find_first_not_of(" \t\r\n") == std::string::npos) continue;
        std::vector<double> row;
        std::stringstream ss(s);
        std::string cell;
        while (std::getline(ss, cell, ',')) {
            char *end = nullptr;
            errno = 0;
            const double x = std::strtod(cell.c_str(), &end);
            row.push_back(end == cell.c_str() ? 0.0 : x);
        }
        rows.push_back(row);
    }
    int32_t nc = 0;
    for (auto &r : rows) nc = std::max<int32_t>(nc, (int32_t)r.size());
    *nrows = (int32_t)rows.size();
    *ncols = nc;
    if (!table) return FX_OK;
    if (*nrows > max_rows || nc > max_cols) return fx_fail(FX_ERR_ARG, "forcing table larger than the buffer");
    for (int32_t i = 0; i < *nrows; i++)
        for (int32_t j = 0; j < max_cols; j++)
            table[(size_t)i * max_cols + j] = j < (int32_t)rows[i].size() ? (float)rows[i][j] : 0.0f;
    return FX_OK;
}

int fx_write_results(const char *path, const fx_results_grids *g)
{
    if (!path || !g || !g->zb || !g->z || !g->h) return fx_fail(FX_ERR_ARG, "path or grids are NULL");
    FILE *f = std::fopen(path, "w");
    if (!f) return fx_fail(FX_ERR_IO, std::string("cannot write ") + path);
    std::fputs("irow [-],icol [-],Xcoord,Ycoord,z [m],h [m],ux [m/s],uy [m/s],qx * dxy [m3/sec],"
               "qy * dxy [m3/sec],us [m/s],conc_SW [mg/l],soil_mass [g],fe_1 [m2/s],fn_1 [m2/s],"
               "twetimetracer [sec]\n", f);
    const int64_t mr = (int64_t)g->nrows + 2;
    // H_MIN_TO_PRINT is held as an unsigned long (GlobVar.h:102): 0.005 -> 0
    const double hmin = g->h_min_print > 0.0 ? (double)(unsigned long)g->h_min_print : 0.0;
    auto val = [&](const double *a, int64_t p) { return a ? a[p] : 0.0; };
    int64_t rows = 0;
    double v[16];
    for (int32_t icol = 1; icol <= g->ncols; icol++) {
        for (int32_t irow = 1; irow <= g->nrows; irow++) {
            const int64_t p = irow + icol * mr;
            if (!(g->zb[p] != 0.0 && g->h[p] > hmin)) continue;
            v[0] = irow;
            v[1] = icol;
            v[2] = (double)icol * g->dxy + g->xllcorner;
            v[3] = (double)irow * g->dxy + g->yllcorner;
            v[4] = g->z[p];
            v[5] = g->z[p] - g->zb[p];
            v[6] = val(g->ux, p);                        // ux, uy: never updated by the solver,
            v[7] = val(g->uy, p);                        // only carried over from a restart file
            v[8] = val(g->qx, p) * g->dxy;
            v[9] = val(g->qy, p) * g->dxy;
            v[10] = val(g->us, p);
            v[11] = val(g->conc0, p);
            v[12] = val(g->soil_mass, p);
            v[13] = val(g->fe1, p);
            v[14] = val(g->fn1, p);
            v[15] = val(g->twet, p);
            for (int k = 0; k < 16; k++) std::fprintf(f, k ? ",%.16e" : "%.16e", v[k]);
            std::fputc('\n', f);
            rows++;
        }
    }
    if (rows == 0) {                                      // the reference saves one (zero) row then
        for (int k = 0; k < 16; k++) std::fprintf(f, k ? ",%.16e" : "%.16e", 0.0);
        std::fputc('\n', f);
    }
    const bool ok = std::fclose(f) == 0;
    return ok ? FX_OK : fx_fail(FX_ERR_IO, std::string("cannot write ") + path);
}

int fx_read_results(const char *path, int64_t max_rows, double *table, int64_t *nrows)
{
    if (!path || !nrows) return fx_fail(FX_ERR_ARG, "path or nrows is NULL");
    std::ifstream f(path);
    if (!f.is_open()) return fx_fail(FX_ERR_IO, std::string("cannot open Results file ") + path);
    std::string s;
    int64_t n = 0;
    bool first = true;
    std::vector<double> row;
    while (std::getline(f, s)) {
        if (first) { first = false; continue; }         // the header line
        if (s.find_first_not_of(" \t\r\n") == std::string::npos) continue;
        row.assign(16, 0.0);
        std::stringstream ss(s);
        std::string cell;
        int j = 0;
        while (j < 16 && std::getline(ss, cell, ',')) {
            char *end = nullptr;
            const double x = std::strtod(cell.c_str(), &end);
            row[j++] = end == cell.c_str() ? 0.0 : x;
        }
        if (table) {
            if (n >= max_rows) return fx_fail(FX_ERR_ARG, "Results file has more rows than the buffer");
            std::copy(row.begin(), row.end(), table + n * 16);
        }
        n++;
    }
    if (first) return fx_fail(FX_ERR_IO, std::string("Results file is empty: ") + path);
    *nrows = n;
    return FX_OK;
}

int fx_find_last_step(const char *folder, int32_t *timstart)
{
    if (!folder || !timstart) return fx_fail(FX_ERR_ARG, "folder or timstart is NULL");
    int t = 0;
    if (DIR *dir = opendir(folder)) {
        while (dirent *e = readdir(dir)) {
            if (!std::strcmp(e->d_name, ".") || !std::strcmp(e->d_name, "..")) continue;
            char head[7] = {0};
            std::strncpy(head, e->d_name, 6);
            t = std::max(t, std::atoi(head));
        }
        closedir(dir);
    }
    *timstart = t;
    return FX_OK;
}

} // extern "C"
