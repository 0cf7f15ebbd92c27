// fluxos_hip_main.cpp -- command-line driver: the reference's main.cpp run
// (modset -> DEM -> forcing -> initiation -> time loop -> Results/<t>.txt)
// over the C ABI of libfluxos_hip.so.
//
//   fluxos_hip <modset.json> [--fp32] [--device N]
//
// Default precision is FX_FP64_PARITY with the exact ADE, which reproduces the
// reference bit for bit; --fp32 selects the FP32-mixed path with the fast ADE.
// RESTART (the newest Results file as initial conditions) and the WINTRA
// soil release are honoured as the reference does them; OPENWQ is refused.
// Steps between two print times run device-resident (fx_advance); the state
// is copied back only to write a Results file.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "fluxos_hip.h"
#include "json_lite.hpp"

namespace {

struct Log {
    std::ofstream f;
    void operator()(const std::string &msg)
    {
        std::cout << msg << std::endl;
        if (f.is_open()) f << msg << "\n";
    }
};

std::string dirname_of(const std::string &p)
{
    const size_t k = p.find_last_of("/\\");
    return k == std::string::npos ? std::string(".") : p.substr(0, k);
}

double num(const fxjson::Value *v, double dflt)
{
    if (!v) return dflt;
    if (v->kind == fxjson::Value::Number) return v->num;
    if (v->kind == fxjson::Value::Bool) return v->b ? 1.0 : 0.0;
    return dflt;
}

bool flag(const fxjson::Value *v) { return v && num(v, 0.0) != 0.0; }

std::string str(const fxjson::Value *v) { return v && v->kind == fxjson::Value::String ? v->str : std::string(); }

int die(Log &log, const std::string &msg, int rc = 1)
{
    log(msg);
    return rc;
}

// the state grids a Results file needs; soil_mass comes from the device only
// while WINTRA changes it, ux / uy never change (initiate.cpp, write_results.cpp:60-61)
struct Grids {
    std::vector<double> zb, z, h, qx, qy, us, conc0, fe1, fn1, twet, soil, ux, uy;
};

int fetch(fx_domain dom, int64_t mm, bool wintra, Grids &g)
{
    auto get = [&](int field, std::vector<double> &v) {
        v.resize(mm);
        return fx_get_field(dom, field, v.data());
    };
    int rc;
    if ((rc = get(FX_F_Z, g.z)) || (rc = get(FX_F_H, g.h)) || (rc = get(FX_F_QX, g.qx)) ||
        (rc = get(FX_F_QY, g.qy)) || (rc = get(FX_F_US, g.us)) || (rc = get(FX_F_CONC0, g.conc0)) ||
        (rc = get(FX_F_FE1, g.fe1)) || (rc = get(FX_F_FN1, g.fn1)) || (rc = get(FX_F_TWET, g.twet)))
        return rc;
    if (wintra && (rc = get(FX_F_SOIL, g.soil))) return rc;
    return FX_OK;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <modset.json> [--fp32] [--device N]\n", argv[0]);
        return 2;
    }
    const std::string modset_path = argv[1];
    bool fp32 = false;
    int device = 0;
    for (int a = 2; a < argc; a++) {
        if (!std::strcmp(argv[a], "--fp32")) fp32 = true;
        else if (!std::strcmp(argv[a], "--device") && a + 1 < argc) device = std::atoi(argv[++a]);
        else {
            std::fprintf(stderr, "unknown argument %s\n", argv[a]);
            return 2;
        }
    }
    Log log;
    log.f.open(dirname_of(modset_path) + "/fluxos_run.log");
    log("FLUXOS (libfluxos_hip: " + std::string(fx_version()) + ")");
    {
        const std::time_t t0 = std::time(nullptr);
        log(std::string("Simulation started... ") + std::ctime(&t0));
    }

    // ---- modset (main.cpp:103-109, read_functions.cpp:26-110) ----------------
    fxjson::Value m;
    {
        std::ifstream f(modset_path);
        if (!f.is_open()) return die(log, "PROBLEM loading of master input file: " + modset_path);
        std::stringstream ss;
        ss << f.rdbuf();
        try {
            m = fxjson::parse(ss.str());
        } catch (const std::exception &e) {
            return die(log, e.what());
        }
    }
    // compulsory keys (read_functions.cpp:39-65): a missing one is the
    // reference's json type_error
    const fxjson::Value *out = m.find("OUTPUT");
    for (const fxjson::Value *v :
         {m.find("DEM_FILE"), m.find("SIM_DATETIME_START"), m.find("RESTART"), out,
          out ? out->find("OUTPUT_FOLDER") : nullptr, out ? out->find("PRINT_STEP") : nullptr,
          out ? out->find("H_MIN_TO_PRINT") : nullptr, m.find("ROUGNESS_HEIGHT"), m.find("SOIL_RELEASE_RATE"),
          m.find("SOIL_CONC_BACKGROUND"), m.at({"EXTERNAL_MODULES", "ADE_TRANSPORT", "STATUS"}),
          m.at({"EXTERNAL_MODULES", "OPENWQ", "STATUS"}), m.at({"EXTERNAL_MODULES", "WINTRA", "STATUS"})})
        if (!v) return die(log, "PROBLEM loading of master input file: " + modset_path);
    const std::string dem_file = str(m.find("DEM_FILE"));
    const std::string out_dir = str(out->find("OUTPUT_FOLDER"));
    const unsigned long print_step = (unsigned long)num(out->find("PRINT_STEP"), 3600.0);
    const double h_min_print = num(out->find("H_MIN_TO_PRINT"), 0.0);
    const double ks_input = num(m.find("ROUGNESS_HEIGHT"), 0.0);
    const double soil_release_rate = num(m.find("SOIL_RELEASE_RATE"), 0.0);
    const double soil_conc_bckgrd = num(m.find("SOIL_CONC_BACKGROUND"), 0.0);
    const bool restart = flag(m.find("RESTART"));
    const bool ade = flag(m.at({"EXTERNAL_MODULES", "ADE_TRANSPORT", "STATUS"}));
    const bool openwq = flag(m.at({"EXTERNAL_MODULES", "OPENWQ", "STATUS"}));
    // without the ADE solver the reference switches WINTRA (and OpenWQ) off (main.cpp:207-213)
    const bool wintra = ade && flag(m.at({"EXTERNAL_MODULES", "WINTRA", "STATUS"}));
    if (openwq && ade) return die(log, "OPENWQ coupling is not available in this build");
    int nchem = 1;
    double dcoef[FX_MAX_CHEM] = {0};
    if (ade) {
        dcoef[0] = num(m.at({"EXTERNAL_MODULES", "ADE_TRANSPORT", "D_COEF"}), 0.0);
        const fxjson::Value *per = m.at({"EXTERNAL_MODULES", "ADE_TRANSPORT", "D_COEF_PER_SOLUTE"});
        if (per && per->kind == fxjson::Value::Array && !per->arr.empty()) {
            nchem = (int)std::min<size_t>(per->arr.size(), FX_MAX_CHEM);
            for (int k = 0; k < nchem; k++) dcoef[k] = num(&per->arr[k], 0.0);
        }
    }
    const std::string meteo_file = str(m.find("METEO_FILE"));
    const fxjson::Value *inflow = m.find("INFLOW_FILE");
    const std::string inflow_file = inflow ? str(inflow->find("FILENAME")) : std::string();
    if (meteo_file.empty() && inflow_file.empty())
        return die(log, "METEO_FILE or INFLOW_FILE must be provided in master file: " + modset_path);
    log("Successful loading of master input file: " + modset_path);

    // ---- DEM (read_functions.cpp:112-225) ------------------------------------
    fx_dem_header hdr{};
    if (fx_read_dem(dem_file.c_str(), &hdr, nullptr)) return die(log, fx_last_error());
    const int nr = hdr.nrows, nc = hdr.ncols;
    const int64_t mr = nr + 2, mm = mr * (nc + 2);
    std::vector<float> dem((size_t)nr * nc);
    if (fx_read_dem(dem_file.c_str(), &hdr, dem.data())) return die(log, fx_last_error());
    log("Successful loading of DEM file: " + dem_file);
    std::vector<double> zb(mm), ks(mm), z(mm), h(mm), conc0(mm);
    std::vector<float> weir(mm), ldry(mm);
    if (fx_read_geo(nr, nc, hdr.nodata, dem.data(), ks_input, zb.data(), ks.data(), weir.data()))
        return die(log, fx_last_error());
    log("Print step (s) = " + std::to_string(print_step));
    log("Roughness height (m) = " + std::to_string(ks_input));
    log("Cell size (m) = " + std::to_string(hdr.cellsize));

    // ---- forcing (read_functions.cpp:237-326) ---------------------------------
    std::vector<float> meteo, inflow_tab;
    int32_t mrows = 0, irows = 0;
    int ntim_meteo = 0, ntim_inflow = 0;
    const int tcols = 2 + nchem;
    // a forcing file as rows of 2+nchem floats (extra columns are dropped)
    auto read_table = [&](const std::string &path, int32_t *nrows, std::vector<float> &tab) {
        int32_t n = 0, w = 0;
        if (fx_read_forcing(path.c_str(), 0, 0, nullptr, &n, &w)) return false;
        const int wide = std::max<int>(w, tcols);
        std::vector<float> all((size_t)std::max(1, n) * wide);
        if (fx_read_forcing(path.c_str(), n, wide, all.data(), &n, &w)) return false;
        tab.assign((size_t)std::max(1, n) * tcols, 0.0f);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < tcols; j++) tab[(size_t)i * tcols + j] = all[(size_t)i * wide + j];
        *nrows = n;
        return true;
    };
    if (!meteo_file.empty()) {
        std::vector<float> rows;
        if (!read_table(meteo_file, &mrows, rows)) return die(log, fx_last_error());
        // row 0 is the header slot the reference leaves empty (read_functions.cpp:258-266)
        meteo.assign((size_t)(mrows + 1) * tcols, 0.0f);
        std::copy(rows.begin(), rows.begin() + (size_t)mrows * tcols, meteo.begin() + tcols);
        if (mrows > 0) ntim_meteo = (int)meteo[(size_t)mrows * tcols];
        log("Successful loading of METEO file: " + meteo_file);
    }
    int64_t in_irow = 0, in_icol = 0;
    if (!inflow_file.empty()) {
        if (!read_table(inflow_file, &irows, inflow_tab)) return die(log, fx_last_error());
        if (irows > 0) ntim_inflow = (int)inflow_tab[(size_t)(irows - 1) * tcols];
        // discharge cell (read_functions.cpp:228-231): coordinates are held as
        // unsigned long and the rounding comes before the division
        const unsigned long xc = (unsigned long)num(inflow->at({"DISCHARGE_LOCATION", "X_COORDINATE"}), 0.0);
        const unsigned long yc = (unsigned long)num(inflow->at({"DISCHARGE_LOCATION", "Y_COORDINATE"}), 0.0);
        if (xc != 0 || yc != 0) {
            in_irow = (int64_t)(unsigned long)(std::round((double)yc - hdr.yllcorner) / hdr.cellsize);
            in_icol = (int64_t)(unsigned long)(std::round((double)xc - hdr.xllcorner) / hdr.cellsize);
        }
        log("Successful loading of INFLOW file: " + inflow_file);
    }
    const int ntim = std::max(ntim_meteo, ntim_inflow);
    log("Simulation time (days) = " + std::to_string(ntim) + " (= " + std::to_string(ntim) + " sec)");

    log("Soil release rate (1/hour) = " + std::to_string(soil_release_rate));
    log("Soil initial background mass available for release to runoff (g) = " + std::to_string(soil_conc_bckgrd));
    if (ade) log(" > ADE_solver activated");
    if (wintra) log(" > WINTRA activated");

    // the output folder is created first (initiate.cpp:40)
    if (!out_dir.empty()) {
        std::error_code ec;
        std::filesystem::create_directories(out_dir, ec);
    }

    // ---- initiation (initiate.cpp:27-231) -----------------------------------------
    std::vector<double> qx0(mm, 0.0), qy0(mm, 0.0), us0(mm, 0.0), twet0(mm, 0.0), soil(mm, 0.0);
    std::vector<double> ux(mm, 0.0), uy(mm, 0.0);
    unsigned timstart = 0;
    bool loaded = false;
    if (restart) {
        // the newest Results file (initiate.cpp:116-120, common.cpp:69-106)
        int32_t ts = 0;
        if (fx_find_last_step((out_dir + "/").c_str(), &ts)) return die(log, fx_last_error());
        timstart = (unsigned)ts;
        const std::string init_file = out_dir + "/" + "/" + std::to_string(ts) + ".txt";
        int64_t n = 0;
        if (fx_read_results(init_file.c_str(), 0, nullptr, &n) == FX_OK) {
            std::vector<double> table((size_t)std::max<int64_t>(1, n) * 16);
            if (fx_read_results(init_file.c_str(), n, table.data(), &n)) return die(log, fx_last_error());
            fx_init_grids ig{zb.data(), z.data(), h.data(), qx0.data(), qy0.data(), us0.data(), ux.data(),
                             uy.data(), conc0.data(), soil.data(), twet0.data(), ldry.data()};
            if (fx_initiation_restart(nr, nc, ade ? 1 : 0, soil_conc_bckgrd, table.data(), n, &ig))
                return die(log, fx_last_error());
            loaded = true;
            log("RESTART option TRUE: Successful loading of initial conditions file: " + init_file);
        } else {
            log("RESTART option TRUE, but NO INITIAL CONDITIONS FOUND: All variables set to zero");
        }
    }
    if (!loaded) {
        if (fx_initiation(nr, nc, ade ? 1 : 0, zb.data(), z.data(), h.data(), ldry.data(), conc0.data()))
            return die(log, fx_last_error());
        if (!restart) log("RESTART option FALSE: All variables set to zero");
    }

    // ---- device domain ----------------------------------------------------------
    fx_domain_desc desc{};
    desc.nrows = nr;
    desc.ncols = nc;
    desc.dxy = hdr.cellsize;
    desc.nchem = nchem;
    desc.hdry = ks[1 + 1 * mr];                     // ds.hdry = ks(1,1) (main.cpp:259)
    desc.nodata = hdr.nodata;
    desc.gacc = 9.80665f;
    desc.cfl = 1.0f;
    desc.cvdef = 0.07f;
    desc.nuem = 1.793e-6f;
    for (int k = 0; k < FX_MAX_CHEM; k++) desc.d_coef[k] = dcoef[k];
    desc.ade_solver = ade ? 1 : 0;
    desc.precision = fp32 ? FX_FP32_MIXED : FX_FP64_PARITY;
    desc.ade_mode = !ade ? FX_ADE_OFF : (fp32 ? FX_ADE_FAST : FX_ADE_EXACT);
    desc.device = device;
    fx_domain dom = nullptr;
    if (fx_domain_create(&desc, &dom)) return die(log, fx_last_error());
    auto fail = [&](const std::string &what) {
        log(what + ": " + fx_last_error());
        fx_domain_destroy(dom);
        return 1;
    };
    if (fx_set_static(dom, zb.data(), ks.data(), weir.data())) return fail("fx_set_static");
    if (fx_set_field(dom, FX_F_Z, z.data()) || fx_set_field(dom, FX_F_H, h.data()) ||
        fx_set_field(dom, FX_F_LDRY, ldry.data()) || fx_set_field(dom, FX_F_CONC0, conc0.data()))
        return fail("fx_set_field");
    if (loaded && (fx_set_field(dom, FX_F_QX, qx0.data()) || fx_set_field(dom, FX_F_QY, qy0.data()) ||
                   fx_set_field(dom, FX_F_US, us0.data()) || fx_set_field(dom, FX_F_TWET, twet0.data())))
        return fail("fx_set_field");
    if (wintra && (fx_set_wintra(dom, 1, soil_release_rate) || fx_set_field(dom, FX_F_SOIL, soil.data())))
        return fail("fx_set_wintra");
    if (!meteo.empty() && fx_set_meteo(dom, mrows + 1, tcols, meteo.data(), nullptr)) return fail("fx_set_meteo");
    if (!inflow_tab.empty() && fx_set_inflow(dom, irows, tcols, inflow_tab.data(), in_irow, in_icol))
        return fail("fx_set_inflow");
    if (fx_set_clock(dom, 0.0, 0.0, 0)) return fail("fx_set_clock");

    // ds.tim is not moved to timstart on a restart (main.cpp:255-261): the
    // clock starts at 0 and the first print is at timstart + PRINT_STEP
    // ---- time loop (main.cpp:271-470) ---------------------------------------------
    unsigned long print_next = timstart + print_step;
    Grids g;
    g.zb.resize(mm);
    g.soil = soil;
    g.ux = ux;
    g.uy = uy;
    if (fx_get_field(dom, FX_F_ZB, g.zb.data())) return fail("fx_get_field");
    auto t_last = std::chrono::steady_clock::now();
    for (;;) {
        fx_step_info info{};
        if (fx_advance(dom, (double)ntim, (double)print_next, 1 << 30, &info)) return fail("fx_advance");
        if (info.status & 1) {
            if (fetch(dom, mm, wintra, g)) return fail("fx_get_field");
            fx_results_grids rg{};
            rg.nrows = nr;
            rg.ncols = nc;
            rg.dxy = hdr.cellsize;
            rg.xllcorner = hdr.xllcorner;
            rg.yllcorner = hdr.yllcorner;
            rg.h_min_print = h_min_print;
            rg.zb = g.zb.data(); rg.z = g.z.data(); rg.h = g.h.data();
            rg.qx = g.qx.data(); rg.qy = g.qy.data(); rg.us = g.us.data();
            rg.conc0 = g.conc0.data(); rg.fe1 = g.fe1.data(); rg.fn1 = g.fn1.data(); rg.twet = g.twet.data();
            rg.soil_mass = g.soil.data(); rg.ux = g.ux.data(); rg.uy = g.uy.data();
            const std::string path = out_dir + "/" + std::to_string((long)std::llround((double)print_next)) + ".txt";
            if (fx_write_results(path.c_str(), &rg)) return fail("Problem when saving the results");
            const auto now = std::chrono::steady_clock::now();
            const double el = std::chrono::duration<double>(now - t_last).count();
            t_last = now;
            log("Saved: '" + std::to_string(print_next) + ".txt' || time step (min): " +
                std::to_string(print_step / 60) + " || Time elapsed (min): " + std::to_string(el / 60));
            print_next += print_step;
        }
        if ((info.status & 2) || info.steps == 0) break;
    }
    log("-----------------------------------------------");
    log("Simulation complete");
    fx_domain_destroy(dom);
    return 0;
}
