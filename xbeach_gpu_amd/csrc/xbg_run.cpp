// Host driver of the surf-beat model over libxbg.so: the reference's main
// program (Wave_gpu.cu:1639-2990) and main loop (mainloopGPU,
// Wave_gpu.cu:610-937) on the C-ABI.
//
//   XBG_param.txt          readparamstr / findparameter (read_input.cpp:595-1474)
//   bathymetry .md/.dep/.bot  readbathyHead / readbathy / readXBbathy
//                          (read_input.cpp:1508-1634); .nc readgridncsize /
//                          readnczb (writenetcdf.cpp:743-927)
//   sea level, wind        readWLfile / readWNDfile (read_input.cpp:259-405)
//   wave boundary          type 1 ReadCstBnd + GenCstWave, type 2 XBeach
//                          files (readXbbnd*), type 3 reuse file
//                          (read_reuse_bndnc*), type 4 ReadJSWPBnd +
//                          makjonswap + GenWGnLBW, type 5 ReadSPECBnd +
//                          readSWANSPEC + GenWGnLBW; generated blocks saved
//                          to XBG_bnd_reuse.nc (read_input.cpp:25-593,
//                          Wavestep.cu:24-382, 512-753)
//   sanity / defaults      checkparamsanity (read_input.cpp:1227-1444)
//   initial state          Wave_gpu.cu:1988-2056, 2629-2691
//   output                 netCDF (creatncfileUD / defncvar / writencvarstep)
//
// Exposed as xbg_run_case() (and the xbg_run executable); the parameter-line
// parser is exposed as xbg_param_line() for tests.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <ctime>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/xbg.h"

namespace {

const double PI_ = 3.14159265;  // read_input.cpp / Wave_gpu.cu (double literal)

std::vector<std::string> split(const std::string& s, char delim) {
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, delim))
        if (!item.empty()) out.push_back(item);
    return out;
}
std::string trim(const std::string& str, const std::string& ws = " ") {
    size_t b = str.find_first_not_of(ws);
    if (b == std::string::npos) return "";
    size_t e = str.find_last_not_of(ws);
    return str.substr(b, e - b + 1);
}
// read_input.cpp:1446-1474: "key = value ; comment"; the key must match
// exactly; the value is the text up to the first ';', trimmed.
std::string findparameter(const std::string& key, const std::string& line) {
    std::vector<std::string> sp = split(line, '=');
    std::string val;
    if (sp.size() > 1) {
        if (trim(sp[0]) == key) {
            std::vector<std::string> r = split(sp[1], ';');
            if (!r.empty()) val = r[0];
        }
    }
    return trim(val);
}
// split a data line by tab, then space, then comma (read_input.cpp:487-505)
std::vector<std::string> fields(const std::string& line, size_t need) {
    std::vector<std::string> e = split(line, '\t');
    if (e.size() < need) e = split(line, ' ');
    if (e.size() < need) e = split(line, ',');
    return e;
}
void copystr(char* dst, size_t n, const std::string& s) {
    std::strncpy(dst, s.c_str(), n - 1);
    dst[n - 1] = 0;
}

struct SL {
    double time, wlev0, wlev1;
};
struct Wind {
    double time, U, V;
};

std::string g_err;
int fail(const std::string& m) {
    g_err = m;
    return XBG_E_ARG;
}

std::string join_dir(const std::string& dir, const std::string& f) {
    if (f.empty() || f[0] == '/' || dir.empty()) return f;
    return dir + "/" + f;
}

// XBG_log.txt of the case (write_text_to_log_file, read_input.cpp:1637-1643:
// one line per message, the file reopened in append mode each time).
struct Log {
    std::string path;
    void line(const std::string& t) const {
        if (path.empty()) return;
        FILE* f = fopen(path.c_str(), "a");
        if (!f) return;
        fprintf(f, "%s\n", t.c_str());
        fclose(f);
    }
};

std::string d2s(double v) { return std::to_string(v); }

// SaveParamtolog (read_input.cpp:1645-1755): the parameters in the
// XBG_param.txt syntax (angles in radians as held internally).
void save_params(const Log& L, const xbg_params& P, const xbg_inputs& IN, const std::vector<std::string>& outvars) {
    auto kv = [&](const std::string& k, const std::string& v, const char* end = ";") { L.line(k + " = " + v + end); };
    L.line("#################################");
    L.line("# Bathymetry file");
    kv("bathy", IN.bathy);
    kv("nx", std::to_string(P.nx));
    kv("ny", std::to_string(P.ny));
    kv("dx", d2s(P.dx));
    kv("grdalpha", d2s(P.grdalpha * 180.0 / PI_));
    L.line("\n");
    L.line("# Model controls");
    kv("swave", std::to_string(P.swave));
    kv("flow", std::to_string(P.flow));
    kv("sedtrans", std::to_string(P.sedtrans));
    kv("morphology", std::to_string(P.morphology));
    kv("gpudevice", "0");
    L.line("\n");
    L.line("# Flow parameters");
    kv("eps", d2s(P.eps));
    kv("cfsand", d2s(P.cfsand));
    kv("cfreef", d2s(P.cfreef));
    kv("usesmago", std::to_string(P.usesmago));
    if (P.usesmago == 0) kv("nuh", d2s(P.nuh));
    else kv("smag", d2s(P.smag));
    kv("nuhfac", d2s(P.nuhfac));
    kv("lat", d2s(P.lat));
    kv("fc", d2s(P.fc));
    kv("Cd", d2s(P.Cd));
    kv("wci", d2s(P.wci));
    kv("hwci", d2s(P.hwci));
    L.line("\n");
    L.line("# Waves parameters");
    kv("breakmodel", std::to_string(P.breakmodel));
    kv("gamma", d2s(P.gammaa));
    kv("n", d2s(P.n));
    kv("alpha", d2s(P.alpha));
    kv("gammax", d2s(P.gammax));
    kv("beta", d2s(P.beta));
    kv("fwsand", d2s(P.fwsand));
    kv("fwreef", d2s(P.fwreef));
    kv("roller", std::to_string(P.roller));
    kv("thetamin", d2s(P.thetamin));
    kv("thetamax", d2s(P.thetamax));
    kv("dtheta", d2s(P.dtheta), "; ");
    kv("ntheta", std::to_string(P.ntheta), "; ");
    L.line("\n");
    L.line("# Wave boundary parameters");
    kv("dtbc", d2s(P.dtbc), "; ");
    kv("rtlength", d2s(P.rtlength), "; ");
    kv("sprdthr", d2s(P.sprdthr), "; ");
    kv("random", std::to_string(P.random), "; ");
    L.line("\n");
    L.line("# Sediment parameters");
    kv("D50", d2s(P.D50));
    kv("D90", d2s(P.D90));
    kv("rhosed", d2s(P.rhosed));
    kv("wws", d2s(P.wws));
    kv("drydzmax", d2s(P.drydzmax));
    kv("wetdzmax", d2s(P.wetdzmax));
    kv("maxslpchg", d2s(P.maxslpchg));
    kv("por", d2s(P.por));
    kv("morfac", d2s(P.morfac));
    kv("sus", d2s(P.sus));
    kv("bed", d2s(P.bed));
    kv("facsk", d2s(P.facsk));
    kv("facas", d2s(P.facas));
    L.line("\n");
    L.line("# Timekeeping parameters");
    kv("CFL", d2s(P.CFL));
    kv("sedstart", d2s(P.sedstart));
    kv("outputtimestep", d2s(P.outputtimestep));
    std::string all;
    for (size_t k = 0; k < outvars.size(); k++) all += (k ? ", " : "") + outvars[k];
    kv("outvars", all);
    kv("endtime", d2s(P.endtime));
    L.line("\n");
    L.line("# Files");
    kv("outfile", IN.outfile);
    kv("SedThkfile", IN.sedthk);
    kv("wavebndfile", IN.wavebndfile);
    kv("wavebndtype", std::to_string(IN.wavebndtype));
    kv("slbndfile", IN.slbndfile);
    kv("windbndfile", IN.windfile);
    for (int o = 0; o < IN.ntsofile && o < IN.ntsnode; o++) {
        kv("TSOfile", IN.tsofile[o]);
        kv("TSnode", std::to_string(IN.tsnode_i[o]) + "," + std::to_string(IN.tsnode_j[o]));
    }
    L.line("\n");
    L.line("# Others");
    kv("g", d2s(P.g));
    kv("rho", d2s(P.rho));
}

}  // namespace

extern "C" {

// readparamstr (read_input.cpp:595-1225) for one line.  Angles thetamin,
// thetamax, dtheta arrive in degrees and are stored in radians; grdalpha
// and lat stay in degrees (converted where the reference converts them).
static int param_line(const char* cline, xbg_params* p, xbg_inputs* in) {
    if (!cline || !p || !in) return XBG_E_ARG;
    std::string line(cline);
    std::string v;
    auto num = [&](const char* k, double& dst) {
        v = findparameter(k, line);
        if (!v.empty()) dst = std::stod(v);
    };
    auto inum = [&](const char* k, int& dst) {
        v = findparameter(k, line);
        if (!v.empty()) dst = std::stoi(v);
    };
    auto str = [&](const char* k, char* dst, size_t n) {
        v = findparameter(k, line);
        if (!v.empty()) copystr(dst, n, v);
    };
    str("bathy", in->bathy, sizeof(in->bathy));
    str("depfile", in->bathy, sizeof(in->bathy));
    inum("swave", p->swave);
    inum("flow", p->flow);
    inum("sedtrans", p->sedtrans);
    inum("morphology", p->morphology);
    inum("roller", p->roller);
    num("eps", p->eps);
    num("cf", p->cf);
    num("cfsand", p->cfsand);
    num("cfreef", p->cfreef);
    num("nuh", p->nuh);
    num("nuhfac", p->nuhfac);
    inum("usesmago", p->usesmago);
    num("smag", p->smag);
    num("lat", p->lat);
    num("Cd", p->Cd);
    num("wci", p->wci);
    num("hwci", p->hwci);
    num("fc", p->hwci);  // the reference stores fc into hwci (read_input.cpp:754-759)
    inum("breakmodel", p->breakmodel);
    num("gamma", p->gammaa);
    num("n", p->n);
    num("alpha", p->alpha);
    num("gammax", p->gammax);
    num("beta", p->beta);
    num("fw", p->fw);
    num("fwsand", p->fwsand);
    num("fwreef", p->fwreef);
    num("D50", p->D50);
    num("D90", p->D90);
    num("rhosed", p->rhosed);
    num("wws", p->wws);
    num("drydzmax", p->drydzmax);
    num("wetdzmax", p->wetdzmax);
    num("maxslpchg", p->maxslpchg);
    num("por", p->por);
    num("morfac", p->morfac);
    num("sus", p->sus);
    num("bed", p->bed);
    num("facsk", p->facsk);
    num("facas", p->facas);
    num("dt", in->dt);
    num("CFL", p->CFL);
    num("sedstart", p->sedstart);
    num("outputtimestep", p->outputtimestep);
    num("endtime", p->endtime);
    str("outfile", in->outfile, sizeof(in->outfile));
    str("SedThkfile", in->sedthk, sizeof(in->sedthk));
    str("wavebndfile", in->wavebndfile, sizeof(in->wavebndfile));
    str("slbndfile", in->slbndfile, sizeof(in->slbndfile));
    str("windbndfile", in->windfile, sizeof(in->windfile));
    v = findparameter("outvars", line);
    if (!v.empty()) {
        // read_input.cpp:1068-1094: keep the supported names, in order
        static const char* sup[] = {"hh", "uu", "vv", "wci", "zs", "zb", "cfm", "dzb", "stdep", "Fx", "Fy",
                                    "cgx", "cgy", "cx", "cy", "ctheta", "D", "E", "H", "urms", "ueu", "vev",
                                    "thetamean", "Hmean", "uumean", "vvmean", "hhmean", "zsmax", "zsmean",
                                    "Cmean", "sigm", "k", "c", "kh", "cg", "sinh2kh", "dhdx", "dhdy", "dudx",
                                    "dudy", "dvdx", "dvdy", "C", "R", "DR", "ee", "vmageu", "vmagev", "dzsdx",
                                    "dzsdy", "dzsdt", "fwm", "hu", "hum", "hv", "hvm", "uv", "vu", "ududx",
                                    "vdvdy", "udvdx", "vdudy", "ust", "rr", "kturb", "rolthick", "ceqsg"};
        std::string acc = in->outvars;
        for (auto& nv : split(v, ',')) {
            std::string vv = trim(nv);
            for (const char* s : sup)
                if (vv == s) {
                    if (!acc.empty()) acc += ",";
                    acc += vv;
                    break;
                }
        }
        copystr(in->outvars, sizeof(in->outvars), acc);
    }
    inum("wavebndtype", in->wavebndtype);
    v = findparameter("thetamin", line);
    if (!v.empty()) p->thetamin = std::stod(v) * PI_ / 180.0;
    v = findparameter("thetamax", line);
    if (!v.empty()) p->thetamax = std::stod(v) * PI_ / 180.0;
    v = findparameter("dtheta", line);
    if (!v.empty()) p->dtheta = std::stod(v) * PI_ / 180.0;
    num("dtbc", p->dtbc);
    num("rtlength", p->rtlength);
    num("sprdthr", p->sprdthr);
    inum("random", p->random);
    inum("nx", p->nx);
    inum("ny", p->ny);
    num("dx", p->dx);
    num("grdalpha", p->grdalpha);
    num("g", p->g);
    num("rho", p->rho);
    num("nmax", p->nmax);
    num("fcutoff", p->fcutoff);
    inum("nspr", p->nspr);
    // read_input.cpp:1018-1040: one output file per TSOfile line, one node
    // per TSnode line, paired in order by checkparamsanity
    v = findparameter("TSOfile", line);
    if (!v.empty()) {
        if (in->ntsofile >= XBG_MAX_TS) return fail("more than " + std::to_string(XBG_MAX_TS) + " TSOfile entries");
        copystr(in->tsofile[in->ntsofile++], sizeof(in->tsofile[0]), v);
    }
    v = findparameter("TSnode", line);
    if (!v.empty()) {
        if (in->ntsnode >= XBG_MAX_TS) return fail("more than " + std::to_string(XBG_MAX_TS) + " TSnode entries");
        std::vector<std::string> e = split(v, ',');
        if (e.size() < 2) return fail("TSnode needs i,j: " + v);
        in->tsnode_i[in->ntsnode] = std::stoi(e[0]);
        in->tsnode_j[in->ntsnode] = std::stoi(e[1]);
        in->ntsnode++;
    }
    return XBG_OK;
}

// readXbbndstep (read_input.cpp:161-220): one XBeach boundary block from
// its binary files of doubles, nrec_read records each (the reference reads
// ceil(rtlength/dtbc)), q in the qfile layout j + k*ny + m*ny*4 and E in
// the Stfile layout j + itheta*ny + m*ny*ntheta; values pass through float
// (DECNUM) as the reference stores them.  Records beyond nrec_read are left
// as they are.
int xbg_xbbnd_read(const char* qpath, const char* epath, int ny, int ntheta, int nrec_read, double* Stfile,
                   double* qfile) {
    if (!qpath || !epath || !Stfile || !qfile || ny < 1 || ntheta < 1 || nrec_read < 1)
        return fail("xbg_xbbnd_read: bad argument");
    auto rd = [](const char* path, size_t n, double* out) -> int {
        FILE* f = fopen(path, "rb");
        if (!f) return fail(std::string("Unable to open file ") + path);
        std::vector<double> v(n);
        size_t got = fread(v.data(), sizeof(double), n, f);
        fclose(f);
        if (got != n)
            return fail(std::string(path) + ": holds " + std::to_string(got) + " values, " + std::to_string(n) +
                        " expected");
        for (size_t k = 0; k < n; k++) out[k] = (double)(float)v[k];
        return XBG_OK;
    };
    if (int rc = rd(qpath, (size_t)4 * ny * nrec_read, qfile)) return rc;
    return rd(epath, (size_t)ntheta * ny * nrec_read, Stfile);
}

void xbg_default_inputs(xbg_inputs* in) {
    std::memset(in, 0, sizeof(*in));
    copystr(in->outfile, sizeof(in->outfile), "XBGPU_output.nc");  // XBeachGPU.h:122
    in->wavebndtype = 2;                                            // XBeachGPU.h:119
}

const char* xbg_run_error(void) { return g_err.c_str(); }

// The reference converts values with std::stod / std::stoi and ends on a
// malformed number; at the C boundary that becomes an error status.
int xbg_param_line(const char* cline, xbg_params* p, xbg_inputs* in) {
    try {
        return param_line(cline, p, in);
    } catch (const std::exception& e) {
        return fail(std::string("malformed parameter line: ") + (cline ? cline : "") + " (" + e.what() + ")");
    }
}

static int run_case(const std::string& pf, const std::string& dir, int verbose, const Log& L) {
    xbg_params P;
    xbg_default_params(&P);
    xbg_inputs IN;
    xbg_default_inputs(&IN);
    {
        std::ifstream fs(pf);
        if (fs.fail()) return fail(pf + " could not be opened");
        std::string line;
        while (std::getline(fs, line))
            if (!line.empty() && line[0] != '#' && xbg_param_line(line.c_str(), &P, &IN)) return XBG_E_ARG;
    }
    // ---- bathymetry (Wave_gpu.cu:1740-1780, read_input.cpp:1508-1634)
    std::string bathy = join_dir(dir, IN.bathy);
    if (bathy.empty()) return fail("No bathymetry file specified");
    std::string ext = split(bathy, '.').back();
    L.line("bathy: " + std::string(IN.bathy));
    L.line("bathy extension: " + ext);
    std::vector<float> zb;
    if (ext == "md") {
        std::ifstream fs(bathy);
        if (fs.fail()) return fail(bathy + " could not be opened");
        std::string line;
        std::getline(fs, line);
        auto e = fields(line, 5);
        if (e.size() < 5) return fail(bathy + ": header needs nx ny dx _ grdalpha");
        P.nx = std::stoi(e[0]);
        P.ny = std::stoi(e[1]);
        P.dx = std::stod(e[2]);
        P.grdalpha = std::stod(e[4]);
        zb.assign((size_t)P.nx * P.ny, 0.0f);
        // rows are written from ny down to 1, each prefixed by its 1-based index
        for (int f = P.ny; f >= 1; f--) {
            int j;
            if (!(fs >> j) || j < 1 || j > P.ny) return fail(bathy + ": bad row index");
            for (int i = 0; i < P.nx; i++)
                if (!(fs >> zb[i + (size_t)(j - 1) * P.nx])) return fail(bathy + ": short row");
        }
    } else if (ext == "dep" || ext == "bot") {
        if (P.nx <= 0 || P.ny <= 0 || P.dx < 1e-7) return fail("nx, ny and dx are required with ." + ext);
        std::ifstream fs(bathy);
        if (fs.fail()) return fail(bathy + " could not be opened");
        zb.assign((size_t)P.nx * P.ny, 0.0f);
        for (size_t k = 0; k < zb.size(); k++)
            if (!(fs >> zb[k])) return fail(bathy + ": too few values");
    } else if (ext == "nc") {
        // readgridncsize / readnczb (writenetcdf.cpp:743-927); grdalpha
        // comes from the parameter file (default 0)
        int gnx = 0, gny = 0;
        double gdx = 0.0;
        if (xbg_nc_grid_size(bathy.c_str(), &gnx, &gny, &gdx)) return fail(xbg_nc_error());
        P.nx = gnx;
        P.ny = gny;
        P.dx = gdx;
        zb.assign((size_t)P.nx * P.ny, 0.0f);
        if (xbg_nc_read_zb(bathy.c_str(), P.nx, P.ny, zb.data())) return fail(xbg_nc_error());
    } else {
        return fail("bathymetry format ." + ext + " is not supported (use .md, .nc, .dep or .bot)");
    }
    P.grdalpha = P.grdalpha * PI_ / 180;  // Wave_gpu.cu:1770
    L.line("nx=" + std::to_string(P.nx) + " ny=" + std::to_string(P.ny) + " dx=" + d2s(P.dx) +
           " grdalpha=" + d2s(P.grdalpha * 180.0 / PI_));
    L.line("Opening sea level bnd...");
    // ---- sea level and wind (Wave_gpu.cu:1783-1830)
    std::vector<SL> slb;
    if (IN.slbndfile[0]) {
        std::ifstream fs(join_dir(dir, IN.slbndfile));
        if (fs.fail()) return fail(std::string(IN.slbndfile) + " could not be opened");
        std::string line;
        while (std::getline(fs, line)) {
            if (line.empty()) continue;
            auto e = fields(line, 2);
            if (e.size() < 2) return fail("sea-level file format");
            SL s{std::stod(e[0]), std::stod(e[1]), std::stod(e[1])};
            if (e.size() >= 3) s.wlev1 = std::stod(e[2]);
            slb.push_back(s);
        }
    } else {
        slb = {{0.0, 0.0, 0.0}, {P.endtime, 0.0, 0.0}};
    }
    L.line("done");
    L.line("Opening wind forcing...");
    std::vector<Wind> wnd;
    if (IN.windfile[0]) {
        std::ifstream fs(join_dir(dir, IN.windfile));
        if (fs.fail()) return fail(std::string(IN.windfile) + " could not be opened");
        std::string line;
        while (std::getline(fs, line)) {
            if (line.empty()) continue;
            auto e = fields(line, 3);
            if (e.size() < 3) return fail("wind file format");
            double spd = std::stod(e[1]), th = (1.5 * PI_ - P.grdalpha) - std::stod(e[2]) * PI_ / 180;
            wnd.push_back({std::stod(e[0]), spd * std::cos(th), spd * std::sin(th)});
        }
    } else {
        double te = std::max(P.endtime, slb.back().time);
        wnd = {{0.0, 0.0, 0.0}, {te, 0.0, 0.0}};
    }
    L.line("done");
    L.line("Reading wave forcing...");
    // ---- wave boundary list (Wave_gpu.cu:1833-1880)
    std::vector<xbg_wavebnd> wb;
    std::vector<std::string> wbq, wbe;  // per record: XBeach q/E files (type 2), spectrum file (type 5)
    const std::string wbfile = join_dir(dir, IN.wavebndfile);
    if (IN.wavebndfile[0] && (IN.wavebndtype == 2 || IN.wavebndtype == 5)) {
        // readXbbndhead / readXbbndfile (read_input.cpp:25-148), ReadSPECBnd
        // (read_input.cpp:534-593): fields split on tabs, else spaces, else
        // commas; a last entry rtlength after the final one
        std::ifstream fs(wbfile);
        if (fs.fail()) return fail(std::string(IN.wavebndfile) + " could not be opened");
        auto split3 = [](const std::string& line, size_t need) {
            std::vector<std::string> e = split(line, '\t');
            if (e.size() < need) e = split(line, ' ');
            if (e.size() < need) e = split(line, ',');
            return e;
        };
        std::string line;
        if (IN.wavebndtype == 2) {
            std::getline(fs, line);
            auto e = split3(line, 5);
            if (e.size() < 5) return fail("Wave bnd file, header format error: 5 fields expected");
            P.thetamin = std::stod(e[0]) * PI_ / 180.0;
            P.thetamax = std::stod(e[1]) * PI_ / 180.0;
            P.dtheta = std::stod(e[2]) * PI_ / 180.0;
            P.dtbc = std::stod(e[3]);
            P.rtlength = std::stod(e[4]);
        }
        const size_t need = IN.wavebndtype == 2 ? 4 : 2;
        while (std::getline(fs, line)) {
            if (line.empty()) continue;
            auto e = split3(line, need);
            if (e.size() < need) return fail("Wave bnd file format error");
            xbg_wavebnd w{};
            w.time = std::stod(e[0]);
            if (IN.wavebndtype == 2) {
                w.Tp = std::stod(e[1]);  // the block's Trep
                wbq.push_back(join_dir(dir, trim(e[2], " \r")));
                wbe.push_back(join_dir(dir, trim(e[3], " \r")));
            } else {
                wbq.push_back(join_dir(dir, trim(e[1], " \r")));
            }
            wb.push_back(w);
        }
        if (wb.empty()) return fail("empty wave boundary file");
        xbg_wavebnd last = wb.back();
        last.time = wb.back().time + P.rtlength;
        wb.push_back(last);
    } else if (IN.wavebndfile[0] && IN.wavebndtype == 3) {
        // read_reuse_bndnc_head / read_reuse_bndnc_vec (writenetcdf.cpp:
        // 1090-1188): bins and block timing from the file, one block per record
        int nrec = 0;
        if (xbg_bndnc_head(wbfile.c_str(), &P, &nrec)) return fail(xbg_nc_error());
        for (int i = 0; i < nrec + 1; i++) {
            xbg_wavebnd w{};
            w.time = P.rtlength * i;
            wb.push_back(w);
        }
    } else if (IN.wavebndfile[0]) {
        std::ifstream fs(wbfile);
        if (fs.fail()) return fail(std::string(IN.wavebndfile) + " could not be opened");
        if (IN.wavebndtype != 1 && IN.wavebndtype != 4)
            return fail("wavebndtype " + std::to_string(IN.wavebndtype) + " is not supported (use 1 to 5)");
        size_t need = IN.wavebndtype == 4 ? 6 : 5;
        std::string line;
        while (std::getline(fs, line)) {
            if (line.empty()) continue;
            auto e = fields(line, need);
            if (e.size() < need) return fail("wave boundary file format");
            xbg_wavebnd w{};
            w.time = std::stod(e[0]);
            w.Hs = std::stod(e[1]);
            w.Tp = std::stod(e[2]);
            w.Dp = (1.5 * PI_ - P.grdalpha) - std::stod(e[3]) * PI_ / 180;
            w.s = std::stod(e[4]);
            w.gamma = need == 6 ? std::stod(e[5]) : 3.3;
            wb.push_back(w);
        }
        if (wb.empty()) return fail("empty wave boundary file");
        if (IN.wavebndtype == 4) {
            xbg_wavebnd last = wb.back();
            last.time = wb.back().time + P.rtlength;  // read_input.cpp:528-529
            wb.push_back(last);
        }
    } else {
        IN.wavebndtype = 1;
        double te = std::max(P.endtime, slb.back().time);
        xbg_wavebnd w{0.0, 0.0001, 10.0, (1.5 * PI_ - P.grdalpha), 1000.0, 3.3};
        wb.push_back(w);
        w.time = te;
        wb.push_back(w);
    }
    // ---- checkparamsanity (read_input.cpp:1227-1444)
    const double tiny = 0.0000001;
    if (P.nx <= 0 || P.ny <= 0 || P.dx < tiny) return fail("nx or ny or dx could not be specified");
    if (std::fabs(P.endtime - 0.0) <= tiny) P.endtime = 1.0 / tiny;
    double endbnd = std::min(slb.back().time, wnd.back().time);
    endbnd = std::min(endbnd, wb.back().time);
    P.endtime = std::min(P.endtime, endbnd);
    if (P.swave == 0) {
        P.ntheta = 1;
        P.dtheta = (P.thetamax - P.thetamin);
    }
    if (std::fabs(P.outputtimestep - 0.0) <= tiny) P.outputtimestep = P.endtime;
    if (std::fabs(P.cfsand - 0.0) < tiny && P.cf < tiny) P.cf = tiny;
    if (std::fabs(P.fwsand - 0.0) < tiny && P.fw < tiny) P.fw = tiny;
    std::vector<std::string> outvars = split(IN.outvars, ',');
    if (outvars.empty() && P.outputtimestep > 0)
        outvars = {"zb", "zs", "uu", "vv", "H", "thetamean", "D", "urms", "ueu", "vev", "C",
                   "dzb", "Fx", "Fy", "hh", "Hmean", "uumean", "vvmean", "hhmean", "zsmean", "Cmean"};
    // time-series outputs paired in order, nodes clamped to the grid
    // (read_input.cpp:1367-1425)
    int nts = std::min(IN.ntsofile, IN.ntsnode);
    if (IN.ntsofile != IN.ntsnode) {
        L.line("WARNING: the number of timeseries output files is not equal to the number of nodes specified");
        if (verbose)
            fprintf(stderr, "WARNING: the number of timeseries output files is not equal to the number of nodes "
                            "specified\n");
    }
    IN.ntsofile = IN.ntsnode = nts;
    for (int o = 0; o < nts; o++) {
        IN.tsnode_i[o] = std::min(std::max(IN.tsnode_i[o], 0), P.nx - 1);
        IN.tsnode_j[o] = std::min(std::max(IN.tsnode_j[o], 0), P.ny - 1);
    }
    xbg_finalize_params(&P);
    // ---- initial state (Wave_gpu.cu:1988-2056)
    L.line("Set initial condition...");
    const int nx = P.nx, ny = P.ny, nt = P.ntheta;
    const size_t N = (size_t)nx * ny;
    std::vector<float> zs(N), hh(N), zero(N, 0.0f), stdep(N, 5.0f);
    for (int f = ny; f >= 1; f--)
        for (int i = 0; i < nx; i++) {
            size_t k = i + (size_t)(f - 1) * nx;
            double zsbnd = slb[0].wlev0 + (slb[0].wlev1 - slb[0].wlev0) * (f - 1) / ny;
            zs[k] = std::max((float)zsbnd, -1 * zb[k]);
            hh[k] = (float)std::max(zb[k] + zsbnd, P.eps);
        }
    const double offdepth = hh[0 + (size_t)std::floor(ny * 0.5) * nx];
    if (IN.sedthk[0]) {
        std::ifstream fs(join_dir(dir, IN.sedthk));
        if (fs.fail()) return fail(std::string(IN.sedthk) + " could not be opened");
        std::string line;
        std::getline(fs, line);
        for (int f = ny; f >= 1; f--) {
            int j;
            if (!(fs >> j) || j < 1 || j > ny) return fail("sediment thickness file: bad row index");
            for (int i = 0; i < nx; i++)
                if (!(fs >> stdep[i + (size_t)(j - 1) * nx])) return fail("sediment thickness file: short row");
        }
    }
    // ---- wave boundary records (waveinitGPU, Wavestep.cu:24-382)
    const int nrec_blk = IN.wavebndtype == 1 ? (int)wb.size() : xbg_bnd_nrec(&P);
    std::vector<double> Stfile((size_t)nrec_blk * nt * ny), qfile((size_t)nrec_blk * 4 * ny);
    double Trep = 10.0;
    std::vector<float> theta(nt), cxs(nt), sxn(nt);
    xbg_theta_tables(&P, theta.data(), cxs.data(), sxn.data());
    // the generated blocks are kept in XBG_bnd_reuse.nc for a later run with
    // wavebndtype 3 (createbndnc / writebndnc, Wavestep.cu:150-199, 560-610)
    const std::string reuse = join_dir(dir, "XBG_bnd_reuse.nc");
    auto save_into = [&](int append, const xbg_wavebnd& w, const std::vector<double>& St,
                         const std::vector<double>& q, double Tr, std::string& err) -> int {
        if (xbg_bndnc_write(reuse.c_str(), append, nrec_blk, ny, nt, P.dx, P.dtheta, P.thetamin, P.dtbc, w.time, w.Hs,
                            Tr, w.Tp, w.Dp, St.data(), q.data())) {
            err = xbg_nc_error();
            return XBG_E_ARG;
        }
        return XBG_OK;
    };
    auto save_block = [&](int append, const xbg_wavebnd& w) -> int {
        std::string e;
        return save_into(append, w, Stfile, qfile, Trep, e) ? fail(e) : XBG_OK;
    };
    // read_reuse_bndnc (writenetcdf.cpp:1190-1286); Trep = 15 as it sets it
    auto read_block = [&](int rec) -> int {
        if (xbg_bndnc_read(wbfile.c_str(), rec, ny, nt, nrec_blk, Stfile.data(), qfile.data()))
            return fail(xbg_nc_error());
        Trep = 15.0;
        return XBG_OK;
    };
    // readXbbndstep (read_input.cpp:161-220): ceil(rtlength/dtbc) records
    auto xb_block = [&](int k) -> int {
        int nread = (int)std::ceil(P.rtlength / P.dtbc);
        if (nread > nrec_blk) nread = nrec_blk;
        if (xbg_xbbnd_read(wbq[k].c_str(), wbe[k].c_str(), ny, nt, nread, Stfile.data(), qfile.data()))
            return XBG_E_ARG;
        Trep = wb[k].Tp;
        return XBG_OK;
    };
    // makjonswap (type 4) or readSWANSPEC (type 5), then GenWGnLBW
    // block k into the given buffers; touches no shared state, so it also
    // runs on the host thread that builds the next block ahead of time
    auto gen_into = [&](int k, std::vector<double>& St, std::vector<double>& q, double& Tr, std::string& err) -> int {
        const xbg_wavebnd& w = wb[k];
        int nf, nd;
        std::vector<double> fr, di, S;
        if (IN.wavebndtype == 5) {
            if (xbg_nc_spectrum_size(wbq[k].c_str(), &nf, &nd)) {
                err = xbg_nc_error();
                return XBG_E_ARG;
            }
            fr.resize(nf);
            di.resize(nd);
            S.resize((size_t)nf * nd);
            if (xbg_nc_spectrum(wbq[k].c_str(), P.grdalpha, fr.data(), di.data(), S.data())) {
                err = xbg_nc_error();
                return XBG_E_ARG;
            }
        } else {
            if (xbg_jonswap_size(&w, &nf, &nd)) {
                err = "jonswap size";
                return XBG_E_ARG;
            }
            fr.resize(nf);
            di.resize(nd);
            S.resize((size_t)nf * nd);
            if (xbg_jonswap(&w, fr.data(), di.data(), S.data())) {
                err = "JONSWAP spectrum: NaN";
                return XBG_E_ARG;
            }
        }
        unsigned seed = 0;
        if (P.random == 1) seed = (unsigned)std::time(nullptr);
        if (xbg_wave_groups(&P, offdepth, nf, nd, fr.data(), di.data(), S.data(), seed, &Tr, St.data(), q.data())) {
            err = "wave-group generation failed";
            return XBG_E_ARG;
        }
        return XBG_OK;
    };
    auto gen_block = [&](int k) -> int {
        std::string e;
        return gen_into(k, Stfile, qfile, Trep, e) ? fail(e) : XBG_OK;
    };
    // Generated boundaries (types 4, 5): block k+1 is built, and appended to
    // the reuse file, on a host thread while the GPU steps through block k
    // (the reference generates it at the block change, Wavestep.cu:545-610).
    struct Ahead {
        std::thread th;
        int k = -1, rc = 0;
        std::string err;
        std::vector<double> St, q;
        double Trep = 0.0;
        ~Ahead() {
            if (th.joinable()) th.join();
        }
    } ahead;
    auto build_ahead = [&](int k) {
        if ((IN.wavebndtype != 4 && IN.wavebndtype != 5) || k < 1 || k >= (int)wb.size() - 1) return;
        if (ahead.th.joinable()) ahead.th.join();
        ahead.k = k;
        ahead.rc = 0;
        ahead.err.clear();
        ahead.St.assign(Stfile.size(), 0.0);
        ahead.q.assign(qfile.size(), 0.0);
        ahead.th = std::thread([&ahead, &gen_into, &save_into, &wb, k]() {
            try {
                ahead.rc = gen_into(k, ahead.St, ahead.q, ahead.Trep, ahead.err);
                if (!ahead.rc) ahead.rc = save_into(1, wb[k], ahead.St, ahead.q, ahead.Trep, ahead.err);
            } catch (const std::exception& e) {
                ahead.err = std::string("wave boundary block: ") + e.what();
                ahead.rc = XBG_E_ARG;
            }
        });
    };
    if (P.swave == 1) {
        if (IN.wavebndtype == 1) {
            std::vector<double> Tp(wb.size());
            if (xbg_constant_waves(&P, (int)wb.size(), wb.data(), theta.data(), Stfile.data(), qfile.data(), Tp.data()))
                return fail("constant wave boundary");
            Trep = Tp[0];
        } else if (IN.wavebndtype == 2) {
            if (xb_block(0)) return XBG_E_ARG;
        } else if (IN.wavebndtype == 3) {
            if (read_block(0)) return XBG_E_ARG;
        } else if (gen_block(0) || save_block(0, wb[0])) {
            return XBG_E_ARG;
        } else {
            build_ahead(1);
        }
    }
    // ---- device state
    xbg_ctx* ctx = nullptr;
    int rc = xbg_create(&P, &ctx);
    if (rc) return fail(std::string("xbg_create: ") + xbg_last_error(nullptr));
    auto chk = [&](int r, const char* what) {
        if (r && g_err.empty()) g_err = std::string(what) + ": " + xbg_last_error(ctx);
        return r;
    };
    std::vector<float> ee(N * nt, 0.0f);
    if (P.swave == 1)
        for (int t = 0; t < nt; t++)
            for (int j = 0; j < ny; j++) ee[(size_t)j * nx + (size_t)t * N] = (float)Stfile[j + (size_t)t * ny];
    rc = chk(xbg_set_field(ctx, "zb", zb.data(), N), "zb") || chk(xbg_set_field(ctx, "zs", zs.data(), N), "zs") ||
         chk(xbg_set_field(ctx, "hh", hh.data(), N), "hh") || chk(xbg_set_field(ctx, "stdep", stdep.data(), N), "stdep") ||
         chk(xbg_set_field(ctx, "ee", ee.data(), N * nt), "ee") || chk(xbg_init_state(ctx, Trep), "init");
    if (!rc && P.swave == 1) rc = chk(xbg_set_bnd_series(ctx, Stfile.data(), qfile.data(), nrec_blk), "bnd series");
    double dtf = 0, dtw = 0, dt = 0;
    if (!rc) rc = chk(xbg_cfl_dt(ctx, 1, &dtf, &dtw), "initial dt");
    dt = dtf * 0.5;  // Wave_gpu.cu:2640
    if (!rc) L.line("Initial timestep: dt=" + d2s(dt));
    if (!rc && dt < 0.00000001) rc = fail("ERROR: timestep too small");
    // ---- output file (Wave_gpu.cu:2940-2959)
    L.line("prepare output");
    xbg_nc* nc = nullptr;
    std::vector<float> buf;
    auto ndim = [&](const std::string& v) {
        if (v == "ee" || v == "rr") return 4;
        if (v == "zb") return P.morphology == 1 ? 3 : 2;
        return 3;
    };
    auto fetch = [&](const std::string& v) -> const float* {
        size_t n = xbg_field_size(ctx, v.c_str());
        if (n == 0) {
            buf.assign(ndim(v) == 4 ? N * nt : N, 0.0f);  // not held on the device: zeros
            return buf.data();
        }
        buf.resize(n);
        if (xbg_get_field(ctx, v.c_str(), buf.data(), n)) return nullptr;
        return buf.data();
    };
    if (!rc && !outvars.empty() && P.outputtimestep > 0) {
        std::string of = join_dir(dir, IN.outfile);
        nc = xbg_nc_create(of.c_str(), nx, ny, nt, P.dx, P.dtheta, P.thetamin, 0.0);
        if (!nc) rc = fail(of + " could not be created");
        for (auto& v : outvars) {
            if (rc) break;
            const float* d = fetch(v);
            if (!d || xbg_nc_def_var(nc, v.c_str(), ndim(v), d)) rc = fail("output variable " + v);
        }
    }
    if (!rc) {
        L.line("...done");
        L.line("#################################");
        L.line("XBeach_GPU Parameters:");
        save_params(L, P, IN, outvars);
        L.line("#################################");
        L.line("Starting Computation ");
    }
    // ---- point time series (Wave_gpu.cu:291-310, 764-781, 875-890): the
    // files start empty; zs and H at the nodes are gathered on the device
    // every step and appended to the files at each output time
    const int ts_cap = 1024;
    std::vector<std::string> ts_path;
    std::vector<double> ts_time;
    std::vector<float> ts_zs, ts_H;
    int ts_ondev = 0;
    if (!rc && nts > 0) {
        std::vector<int> cells(nts);
        for (int o = 0; o < nts; o++) {
            cells[o] = IN.tsnode_i[o] + IN.tsnode_j[o] * nx;
            ts_path.push_back(join_dir(dir, IN.tsofile[o]));
            FILE* f = fopen(ts_path.back().c_str(), "w");
            if (!f) {
                rc = fail(ts_path.back() + " could not be created");
                break;
            }
            fclose(f);
        }
        if (!rc) rc = chk(xbg_set_points(ctx, cells.data(), nts, ts_cap), "points");
    }
    auto ts_drain = [&]() -> int {
        int n = 0;
        std::vector<float> z((size_t)ts_cap * nts), h((size_t)ts_cap * nts);
        if (int r = chk(xbg_read_points(ctx, z.data(), h.data(), ts_cap, &n), "points")) return r;
        ts_zs.insert(ts_zs.end(), z.begin(), z.begin() + (size_t)n * nts);
        ts_H.insert(ts_H.end(), h.begin(), h.begin() + (size_t)n * nts);
        ts_ondev = 0;
        return XBG_OK;
    };
    auto ts_write = [&]() -> int {
        if (nts == 0) return XBG_OK;
        if (int r = ts_drain()) return r;
        for (int o = 0; o < nts; o++) {
            FILE* f = fopen(ts_path[o].c_str(), "a");
            if (!f) return fail(ts_path[o] + " could not be opened");
            for (size_t n = 0; n < ts_time.size(); n++)
                fprintf(f, "%f\t%.4f\t%.4f\n", ts_time[n], (double)ts_zs[n * nts + o], (double)ts_H[n * nts + o]);
            fclose(f);
        }
        ts_time.clear();
        ts_zs.clear();
        ts_H.clear();
        return XBG_OK;
    };
    double totaltime = 0.0, nextoutputtime = P.outputtimestep;
    int WAV = 1, SLs = 1, WNs = 1, nstep = 0;
    if (!rc) rc = chk(xbg_reset_means(ctx), "means");
    // ---- main loop (Wave_gpu.cu:610-937)
    while (!rc && totaltime < P.endtime) {
        nstep++;
        double wdt = dt;
        if ((rc = chk(xbg_cfl_dt(ctx, 0, &dtf, &dtw), "cfl"))) break;
        rc = xbg_next_dt(dtf, dtw, P.swave == 1, wdt, totaltime, nextoutputtime, P.endtime, &dt);
        if (rc) {
            g_err = "Model crashed!";
            L.line("Model crashed!. Exiting.");
            break;
        }
        totaltime = totaltime + dt;
        // wavebnd (Wavestep.cu:512-753): record list / block bookkeeping
        int rec = 0;
        double tnext = 1.0, tsl = 0.0;
        if (P.swave == 1) {
            if (wb[WAV].time - totaltime < 0.0) {
                WAV++;
                if (WAV >= (int)wb.size()) WAV = (int)wb.size() - 1;
                if (IN.wavebndtype >= 2) {
                    int r;
                    const bool ready = ahead.k == WAV - 1;
                    if (ahead.th.joinable()) ahead.th.join();  // a block built for another index is dropped
                    ahead.k = -1;
                    if (ready) {
                        r = ahead.rc ? fail(ahead.err) : XBG_OK;
                        if (!r) {
                            Stfile.swap(ahead.St);
                            qfile.swap(ahead.q);
                            Trep = ahead.Trep;
                        }
                    } else {
                        r = IN.wavebndtype == 2   ? xb_block(WAV - 1)
                            : IN.wavebndtype == 3 ? read_block(WAV - 1)
                                                  : (gen_block(WAV - 1) || save_block(1, wb[WAV - 1]));
                    }
                    if (!r) build_ahead(WAV);
                    if (r) {
                        rc = XBG_E_ARG;
                        break;
                    }
                    if ((rc = chk(xbg_set_bnd_series(ctx, Stfile.data(), qfile.data(), nrec_blk), "bnd series"))) break;
                }
            }
            if (IN.wavebndtype == 1) {
                rec = WAV - 1;
                tnext = wb[WAV].time - wb[WAV - 1].time;
                tsl = totaltime - wb[WAV - 1].time;
            } else {
                rec = (int)std::min(std::floor((totaltime - wb[WAV - 1].time) / P.dtbc),
                                    std::ceil(P.rtlength / P.dtbc) - 1);
                tnext = P.dtbc;
                tsl = totaltime - (rec * P.dtbc + wb[WAV - 1].time);
            }
        }
        // flowbnd (Wave_gpu.cu:1020-1103): sea level, then wind
        if (slb[SLs].time - totaltime < 0.0 && SLs + 1 < (int)slb.size()) SLs++;
        double zsbndi = xbg_interptime(slb[SLs].wlev0, slb[SLs - 1].wlev0, slb[SLs].time - slb[SLs - 1].time,
                                       totaltime - slb[SLs - 1].time);
        double zsbndn = xbg_interptime(slb[SLs].wlev1, slb[SLs - 1].wlev1, slb[SLs].time - slb[SLs - 1].time,
                                       totaltime - slb[SLs - 1].time);
        if (wnd[WNs].time - totaltime < 0.0 && WNs + 1 < (int)wnd.size()) WNs++;
        double windu = xbg_interptime(wnd[WNs].U, wnd[WNs - 1].U, wnd[WNs].time - wnd[WNs - 1].time,
                                      totaltime - wnd[WNs - 1].time);
        double windv = xbg_interptime(wnd[WNs].V, wnd[WNs - 1].V, wnd[WNs].time - wnd[WNs - 1].time,
                                      totaltime - wnd[WNs - 1].time);
        if (P.swave == 1) {
            if ((rc = chk(xbg_use_bnd_record(ctx, rec, tsl, tnext, Trep, zsbndi, zsbndn, windu, windv), "bnd"))) break;
        } else {
            if ((rc = chk(xbg_set_flow_forcing(ctx, nullptr, nullptr, tsl, tnext, zsbndi, zsbndn, windu, windv),
                          "flow forcing")))
                break;
        }
        if ((rc = chk(xbg_step(ctx, dt, totaltime), "step"))) break;
        if ((rc = chk(xbg_accumulate_means(ctx), "means"))) break;
        if (nts > 0) {
            if (ts_ondev == ts_cap && (rc = ts_drain())) break;
            if ((rc = chk(xbg_record_points(ctx), "points"))) break;
            ts_ondev++;
            ts_time.push_back(totaltime);
        }
        if (nextoutputtime - totaltime <= dt * 0.001 && P.outputtimestep > 0) {
            L.line("Writing outputs, totaltime: " + d2s(totaltime) + ", Mean dt= " + d2s(P.outputtimestep / nstep));
            nextoutputtime = nextoutputtime + P.outputtimestep;
            if ((rc = chk(xbg_divide_means(ctx, nstep), "means"))) break;
            if (nc) {
                if (xbg_nc_write_time(nc, totaltime)) rc = fail("output time");
                for (auto& v : outvars) {
                    if (rc || ndim(v) == 2) continue;
                    const float* d = fetch(v);
                    if (!d || xbg_nc_write_var_step(nc, v.c_str(), d)) rc = fail("output " + v);
                }
            }
            if (!rc) rc = ts_write();
            if (verbose) printf("Writing output, totaltime:%f s, Mean dt=%f\n", totaltime, P.outputtimestep / nstep);
            if (!rc) rc = chk(xbg_reset_means(ctx), "means");
            nstep = 0;
        }
    }
    if (nc) xbg_nc_close(nc);
    xbg_destroy(ctx);
    return rc;
}

int xbg_run_case(const char* paramfile, int verbose) {
    if (!paramfile) return fail("NULL parameter file");
    g_err.clear();
    std::string pf(paramfile), dir;
    size_t sl = pf.find_last_of('/');
    if (sl != std::string::npos) dir = pf.substr(0, sl);
    // XBG_log.txt beside the parameter file, started afresh (Wave_gpu.cu:1670-1690)
    Log L{join_dir(dir, "XBG_log.txt")};
    if (FILE* f = fopen(L.path.c_str(), "w")) fclose(f);
    else L.path.clear();
    std::time_t now = std::time(nullptr);
    char tbuf[64];
    std::strftime(tbuf, sizeof(tbuf), "%c", std::localtime(&now));
    L.line("#################################");
    L.line("XBeach_GPU (libxbg, HIP)");
    L.line("#################################");
    L.line(std::string("model started at ") + tbuf);
    const std::clock_t c0 = std::clock();
    int rc;
    try {
        rc = run_case(pf, dir, verbose, L);
    } catch (const std::exception& e) {
        // a malformed number in an input file (std::stod / std::stoi)
        rc = fail(std::string("input error: ") + e.what());
    }
    if (rc) {
        L.line("ERROR: " + g_err);
        if (verbose) fprintf(stderr, "%s\n", g_err.c_str());
    } else {
        L.line("End Computation");
        L.line("#################################");
        L.line("Total runtime= " + std::to_string((double)(std::clock() - c0) / CLOCKS_PER_SEC) + " seconds");
    }
    return rc;
}

}  // extern "C"
