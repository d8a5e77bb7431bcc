// Drop-in file formats of DENISE-SH, host side:
//   parameter file  - positional ".inp" (read_par.c:82-505: the N-th
//                     non-comment line after the first line is parameter N)
//                     or line-based JSON (json_parser.c, read_par_json.c);
//   workflow table  - header line + one row of 24 numbers per stage
//                     (read_par_inv.c:31-34, stage count denise.c:208-214);
//   source file     - count, then "x z y tshift fc amp azimuth type"
//                     (sources.c:36-77);
//   receiver file   - "x y" per line, grid index iround((x+REFREC)/DH),
//                     duplicates dropped (receiver.c:28-74);
//   model files     - raw float32, x-major (all NY values of column 1 first)
//                     (readmod.c:82-104, writemod.c, mergemod.c);
//   SU traces       - 240-byte SEG-Y trace header + ns float32 samples, no
//                     file header (outseis_glob.c:46-265, inseis.c:71-84).
#include "dn_core.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

struct dn_par {
    std::map<std::string, std::string> kv;
    std::string err;
};

struct dn_workflow {
    std::vector<std::vector<float>> rows;
};

namespace {

// read_par.c:82-505: (name, kind) of the parameter on the N-th line.
// kinds: i int, f float, s string, F list of L floats, "ff" two floats,
// "iii"/"iiii" lists of ints.
struct ParDef {
    const char *name;
    const char *kind;
};
const ParDef kInp[] = {
    {"NPROCX", "i"}, {"NPROCY", "i"}, {"FDORDER", "i"}, {"MAXRELERROR", "i"}, {"NX", "i"}, {"NY", "i"},
    {"DH", "f"}, {"TIME", "f"}, {"DT", "f"}, {"QUELLART", "i"}, {"QUELLTYP", "i"}, {"ANGLE", "f"},
    {"PLANE_WAVE_DEPTH", "f"}, {"PHI", "f"}, {"SIGNAL_FILE", "s"}, {"TS", "f"}, {"SRCREC", "i"},
    {"SOURCE_FILE", "s"}, {"RUN_MULTIPLE_SHOTS", "i"}, {"READMOD", "i"}, {"MFILE", "s"}, {"L", "i"},
    {"FL", "F"}, {"TAU", "f"}, {"FREE_SURF", "i"}, {"FW", "i"}, {"DAMPING", "f"}, {"FPML", "f"},
    {"BOUNDARY", "i"}, {"SNAP", "i"}, {"TSNAP1", "f"}, {"TSNAP2", "f"}, {"TSNAPINC", "f"}, {"IDX", "i"},
    {"IDY", "i"}, {"SNAP_FORMAT", "i"}, {"SNAP_FILE", "s"}, {"SEISMO", "i"}, {"READREC", "i"},
    {"REC_FILE", "s"}, {"REFREC", "ff"}, {"XREC1,YREC1", "ff"}, {"XREC2,YREC2", "ff"}, {"NGEOPH", "i"},
    {"REC_ARRAY", "i"}, {"REC_ARRAY_DEPTH", "f"}, {"REC_ARRAY_DIST", "f"}, {"DRX", "i"}, {"NDT", "i"},
    {"SEIS_FORMAT", "i"}, {"SEIS_FILE_VX", "s"}, {"SEIS_FILE_VY", "s"}, {"SEIS_FILE_CURL", "s"},
    {"SEIS_FILE_DIV", "s"}, {"SEIS_FILE_P", "s"}, {"LOG_FILE", "s"}, {"LOG", "i"}, {"CHECKPTREAD", "i"},
    {"CHECKPTWRITE", "i"}, {"CHECKPTFILE", "s"}, {"ITERMAX", "i"}, {"JACOBIAN", "s"}, {"DATA_DIR", "s"},
    {"TAPER", "i"}, {"TAPERLENGTH", "i"}, {"INVTYPE", "i"}, {"GRADT1,GRADT2,GRADT3,GRADT4", "iiii"},
    {"INVMAT1", "i"}, {"INVMAT", "i"}, {"QUELLTYPB", "i"}, {"TESTSHOT_START,TESTSHOT_END,TESTSHOT_INCR", "iii"},
    {"SWS_TAPER_GRAD_VERT", "i"}, {"SWS_TAPER_GRAD_HOR", "i"}, {"SWS_TAPER_GRAD_SOURCES", "i"},
    {"SWS_TAPER_CIRCULAR_PER_SHOT", "i"}, {"SRTSHAPE", "i"}, {"SRTRADIUS", "f"}, {"FILTSIZE", "i"},
    {"SWS_TAPER_FILE", "i"}, {"SPATFILTER", "i"}, {"WD_DAMP", "f"}, {"SPAT_FILT_SIZE", "i"},
    {"SPAT_FILT_1", "i"}, {"SPAT_FILT_ITER", "i"}, {"INV_RHO_ITER", "i"}, {"INV_MODELFILE", "s"},
    {"nfstart", "i"}, {"nf", "i"}, {"nfstart_jac", "i"}, {"nf_jac", "i"}, {"VPUPPERLIM", "f"},
    {"VPLOWERLIM", "f"}, {"VSUPPERLIM", "f"}, {"VSLOWERLIM", "f"}, {"RHOUPPERLIM", "f"}, {"RHOLOWERLIM", "f"},
    {"npower", "f"}, {"k_max_PML", "f"}, {"HESSIAN", "i"}, {"FC_HESS_START", "f"}, {"FC_HESS_INC", "f"},
    {"NFREQ", "i"}, {"FC_HESSIAN", "f"}, {"ORDER_HESSIAN", "i"}, {"GRAD_METHOD", "i"}, {"MODEL_FILTER", "i"},
    {"FILT_SIZE", "i"}, {"INV_STF", "i"}, {"PARA", "s"}, {"N_STF", "i"}, {"N_STF_START", "i"},
    {"TIME_FILT", "i"}, {"FC_START", "f"}, {"FC_END", "f"}, {"FC_INCR", "f"}, {"ORDER", "i"}, {"LNORM", "i"},
    {"DTINV", "i"}, {"EPS_SCALE", "f"}, {"STEPMAX", "i"}, {"SCALEFAC", "f"}, {"PRO", "f"}, {"TRKILL", "i"},
    {"TRKILL_FILE", "s"}, {"TIMEWIN", "i"}, {"PICKS_FILE", "s"}, {"TWLENGTH_PLUS", "f"},
    {"TWLENGTH_MINUS", "f"}, {"GAMMA", "f"}, {"NORMALIZE", "i"}, {"MISFIT_LOG_FILE", "s"},
    {"INV_VP_ITER", "i"}, {"INV_VS_ITER", "i"}, {"MIN_ITER", "i"}, {"GRAD_FILTER", "i"},
    {"FILT_SIZE_GRAD", "i"}, {"TIMELAPSE", "i"}, {"DATA_DIR_T0", "s"},
};
const int kNInp = (int)(sizeof(kInp) / sizeof(kInp[0]));

// read_par_inv.c:32: columns of one workflow row
const char *kWorkflowCols[24] = {"PRO", "TIME_FILT", "FC_START", "FC_END", "ORDER", "TIMEWIN", "GAMMA",
                                 "TWLENGTH_MINUS", "TWLENGTH_PLUS", "INV_VP_ITER", "INV_VS_ITER", "INV_RHO_ITER",
                                 "SPATFILTER", "WD_DAMP1", "WD_DAMP", "EPRECOND", "LNORM", "INV_STF",
                                 "OFFSETC_STF", "EPS_STF", "NORMALIZE", "OFFSET_MUTE", "OFFSETC", "SCALERHO"};

std::vector<std::string> split_ws(const std::string &s)
{
    std::vector<std::string> out;
    std::string cur;
    for (char ch : s) {
        if (std::isspace((unsigned char)ch) || ch == ',') {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else
            cur += ch;
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

std::vector<std::string> split_names(const char *s)
{
    std::vector<std::string> out;
    std::string cur;
    for (const char *p = s; *p; ++p) {
        if (*p == ',') {
            out.push_back(cur);
            cur.clear();
        } else
            cur += *p;
    }
    out.push_back(cur);
    return out;
}

bool read_inp(const char *path, dn_par *P)
{
    std::ifstream f(path);
    if (!f) {
        P->err = std::string("cannot open ") + path;
        return false;
    }
    std::string line;
    int lineno = 0;
    bool first = true;
    std::vector<std::string> lines;
    while (std::getline(f, line)) {
        if (first) {  // the reference only starts counting after the first newline
            first = false;
            continue;
        }
        if (!line.empty() && line[0] == '#') continue;
        lines.push_back(line);
    }
    int L = 0;
    for (const std::string &ln : lines) {
        if (lineno >= kNInp) break;
        const ParDef &d = kInp[lineno++];
        // fscanf "%s =": the label is the first whitespace-delimited token (it
        // may itself contain '='), then optional blanks and the '='
        size_t p = 0;
        while (p < ln.size() && std::isspace((unsigned char)ln[p])) ++p;
        while (p < ln.size() && !std::isspace((unsigned char)ln[p])) ++p;
        while (p < ln.size() && std::isspace((unsigned char)ln[p])) ++p;
        const std::string val = (p < ln.size() && ln[p] == '=') ? ln.substr(p + 1) : std::string();
        std::vector<std::string> tok = split_ws(val);
        const std::string kind = d.kind;
        if (kind == "s" || kind == "i" || kind == "f") {
            P->kv[d.name] = tok.empty() ? "" : tok[0];
            if (std::string(d.name) == "L" && !tok.empty()) L = std::atoi(tok[0].c_str());
            if (std::string(d.name) == "FW" && !tok.empty() && std::atoi(tok[0].c_str()) < 0) P->kv["FW"] = "0";
        } else if (kind == "F") {  // FL: L values
            std::string joined;
            for (int k = 0; k < L && k < (int)tok.size(); ++k) joined += (k ? " " : "") + tok[k];
            P->kv[d.name] = joined;
        } else {  // several named values
            std::vector<std::string> names = split_names(d.name);
            if (std::string(d.name) == "REFREC") names = {"REFRECX", "REFRECY"};
            for (size_t k = 0; k < names.size(); ++k) P->kv[names[k]] = k < tok.size() ? tok[k] : "0";
        }
    }
    if (lineno < kNInp) {
        P->err = "parameter file has " + std::to_string(lineno) + " parameter lines, expected " +
                 std::to_string(kNInp);
        return false;
    }
    return true;
}

// json_parser.c:35-150: lines with a colon, a comma and double quotes and
// without "comment"; each holds one or more  "A,B" : "1,2"  pairs.
bool read_json(const char *path, dn_par *P)
{
    std::ifstream f(path);
    if (!f) {
        P->err = std::string("cannot open ") + path;
        return false;
    }
    std::string line;
    while (std::getline(f, line)) {
        if (line.find(':') == std::string::npos || line.find(',') == std::string::npos ||
            line.find('"') == std::string::npos)
            continue;
        if (line.find("comment") != std::string::npos || line.find("Comment") != std::string::npos) continue;
        std::vector<size_t> q;
        for (size_t k = 0; k < line.size(); ++k)
            if (line[k] == '"') q.push_back(k);
        if (q.size() != 4) continue;
        const std::string names = line.substr(q[0] + 1, q[1] - q[0] - 1);
        const std::string vals = line.substr(q[2] + 1, q[3] - q[2] - 1);
        std::vector<std::string> n = split_names(names.c_str()), v = split_names(vals.c_str());
        for (size_t k = 0; k < n.size() && k < v.size(); ++k) {
            std::string key = n[k], val = v[k];
            while (!key.empty() && std::isspace((unsigned char)key.front())) key.erase(key.begin());
            while (!key.empty() && std::isspace((unsigned char)key.back())) key.pop_back();
            while (!val.empty() && std::isspace((unsigned char)val.front())) val.erase(val.begin());
            while (!val.empty() && std::isspace((unsigned char)val.back())) val.pop_back();
            P->kv[key] = val;
        }
    }
    // FL1..FL4 -> FL (read_par_json.c:308-331)
    if (P->kv.count("L")) {
        const int L = std::atoi(P->kv["L"].c_str());
        std::string joined;
        for (int k = 1; k <= L; ++k) {
            auto it = P->kv.find("FL" + std::to_string(k));
            if (it != P->kv.end()) joined += (joined.empty() ? "" : " ") + it->second;
        }
        if (!joined.empty()) P->kv["FL"] = joined;
    }
    if (P->kv.count("FW") && std::atoi(P->kv["FW"].c_str()) < 0) P->kv["FW"] = "0";
    return true;
}

}  // namespace

extern "C" dn_par *dn_par_read(const char *path)
{
    if (!path) return nullptr;
    dn_par *P = new dn_par();
    const bool ok = std::strstr(path, ".json") ? read_json(path, P) : read_inp(path, P);
    if (!ok) {
        dn_fail(nullptr, "%s", P->err.c_str());
        delete P;
        return nullptr;
    }
    return P;
}

extern "C" void dn_par_free(dn_par *P) { delete P; }

extern "C" int dn_par_get_string(const dn_par *P, const char *key, char *buf, int n)
{
    if (!P || !key || !buf || n < 1) return -1;
    auto it = P->kv.find(key);
    if (it == P->kv.end()) return -1;
    std::snprintf(buf, (size_t)n, "%s", it->second.c_str());
    return 0;
}

extern "C" int dn_par_get_int(const dn_par *P, const char *key, int *out)
{
    if (!P || !key || !out) return -1;
    auto it = P->kv.find(key);
    if (it == P->kv.end() || it->second.empty()) return -1;
    *out = std::atoi(it->second.c_str());
    return 0;
}

extern "C" int dn_par_get_float(const dn_par *P, const char *key, float *out)
{
    if (!P || !key || !out) return -1;
    auto it = P->kv.find(key);
    if (it == P->kv.end() || it->second.empty()) return -1;
    *out = std::strtof(it->second.c_str(), nullptr);
    return 0;
}

extern "C" int dn_par_get_floats(const dn_par *P, const char *key, float *out, int max)
{
    if (!P || !key || !out) return -1;
    auto it = P->kv.find(key);
    if (it == P->kv.end()) return -1;
    std::vector<std::string> tok = split_ws(it->second);
    int n = 0;
    for (; n < (int)tok.size() && n < max; ++n) out[n] = std::strtof(tok[n].c_str(), nullptr);
    return n;
}

extern "C" dn_workflow *dn_workflow_read(const char *path)
{
    std::ifstream f(path ? path : "");
    if (!f) {
        dn_fail(nullptr, "cannot open workflow file %s", path ? path : "(null)");
        return nullptr;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    int nl = 0;
    for (char ch : text)
        if (ch == '\n') ++nl;
    const int stagemax = nl - 1;  // denise.c:208-214
    std::istringstream in(text);
    std::string tok;
    for (int k = 0; k < 24; ++k) in >> tok;  // header
    dn_workflow *w = new dn_workflow();
    for (int s = 0; s < stagemax; ++s) {
        std::vector<float> row(24, 0.0f);
        for (int k = 0; k < 24; ++k) {
            if (!(in >> tok)) {
                dn_fail(nullptr, "workflow stage %d: expected 24 columns", s + 1);
                delete w;
                return nullptr;
            }
            row[k] = std::strtof(tok.c_str(), nullptr);
        }
        w->rows.push_back(row);
    }
    return w;
}

extern "C" void dn_workflow_free(dn_workflow *w) { delete w; }
extern "C" int dn_workflow_nstages(const dn_workflow *w) { return w ? (int)w->rows.size() : -1; }

extern "C" int dn_workflow_get(const dn_workflow *w, int stage, const char *column, float *out)
{
    if (!w || !column || !out || stage < 1 || stage > (int)w->rows.size()) return -1;
    for (int k = 0; k < 24; ++k)
        if (std::strcmp(kWorkflowCols[k], column) == 0) {
            *out = w->rows[stage - 1][k];
            return 0;
        }
    return -1;
}

static int iround(float x) { return (int)std::floor((double)x + 0.5); }
static int iroundd(double x) { return (int)std::floor(x + 0.5); }  // fd.h:18 on a double expression

extern "C" int dn_read_sources(const char *path, int max, float *srcpos, int *nsrc)
{
    FILE *fp = std::fopen(path ? path : "", "r");
    if (!fp) return dn_fail(nullptr, "Source file could not be opened: %s", path ? path : "(null)");
    int n = 0;
    if (std::fscanf(fp, "%d", &n) != 1 || n < 0) {
        std::fclose(fp);
        return dn_fail(nullptr, "source file %s: missing count", path);
    }
    std::rewind(fp);
    char line[256];
    if (!std::fgets(line, 255, fp)) n = 0;
    int l = 0;
    for (; l < n; ++l) {
        if (!std::fgets(line, 255, fp)) break;
        float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const int nv = std::sscanf(line, "%f%f%f%f%f%f%f%f", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7]);
        if (nv < 6) {
            std::fclose(fp);
            return dn_fail(nullptr, "Missing parameter in SOURCE_FILE (source %d)", l + 1);
        }
        if (l < max) {
            float *row = srcpos + (size_t)l * 8;   // [x, y, 0, tshift, fc, amp, azimuth, type]
            row[0] = v[0];
            row[1] = v[2];
            row[2] = 0.0f;
            row[3] = v[3];
            row[4] = v[4];
            row[5] = v[5];
            row[6] = nv > 6 ? v[6] : 0.0f;
            row[7] = nv > 7 ? v[7] : 0.0f;
        }
    }
    std::fclose(fp);
    *nsrc = l;
    return 0;
}

extern "C" int dn_read_receivers(const char *path, float dh, float refrec_x, float refrec_y, int max, int *rec_i,
                                 int *rec_j, int *ntr)
{
    FILE *fp = std::fopen(path ? path : "", "r");
    if (!fp) return dn_fail(nullptr, "Receiver file could not be opened: %s", path ? path : "(null)");
    int n = 0, ch;
    while ((ch = std::fgetc(fp)) != EOF)
        if (ch == '\n') ++n;
    std::rewind(fp);
    std::vector<int> xi(n), yi(n);
    for (int t = 0; t < n; ++t) {
        float x = 0, y = 0;
        if (std::fscanf(fp, "%f%f\n", &x, &y) != 2) {
            n = t;
            break;
        }
        xi[t] = iround((x + refrec_x) / dh);
        yi[t] = iround((y + refrec_y) / dh);
    }
    std::fclose(fp);
    // receiver.c:50-63: later duplicates are dropped, as are x indices <= 0
    std::vector<char> drop(n, 0);
    for (int a = 0; a < n; ++a)
        for (int b = a + 1; b < n; ++b)
            if (!drop[a] && xi[a] == xi[b] && yi[a] == yi[b]) drop[b] = 1;
    int k = 0;
    for (int t = 0; t < n; ++t)
        if (!drop[t] && xi[t] > 0) {
            if (k < max) {
                rec_i[k] = xi[t];
                rec_j[k] = yi[t];
            }
            ++k;
        }
    *ntr = k;
    return 0;
}

// x-major float32 <-> row-major [ny][nx]
extern "C" int dn_read_model_file(const char *path, int nx, int ny, float *out)
{
    FILE *fp = std::fopen(path ? path : "", "rb");
    if (!fp) return dn_fail(nullptr, "Could not open model file %s", path ? path : "(null)");
    std::vector<float> col((size_t)ny);
    for (int i = 0; i < nx; ++i) {
        if (std::fread(col.data(), sizeof(float), (size_t)ny, fp) != (size_t)ny) {
            std::fclose(fp);
            return dn_fail(nullptr, "model file %s is shorter than %d x %d", path, nx, ny);
        }
        for (int j = 0; j < ny; ++j) out[(size_t)j * nx + i] = col[j];
    }
    std::fclose(fp);
    return 0;
}

extern "C" int dn_write_model_file(const char *path, int nx, int ny, const float *in)
{
    FILE *fp = std::fopen(path ? path : "", "wb");
    if (!fp) return dn_fail(nullptr, "Could not open %s for writing", path ? path : "(null)");
    std::vector<float> col((size_t)ny);
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) col[j] = in[(size_t)j * nx + i];
        std::fwrite(col.data(), sizeof(float), (size_t)ny, fp);
    }
    std::fclose(fp);
    return 0;
}

// readmod.c:38-104 (INVMAT1 1): MFILE.vs, MFILE.rho, MFILE.qs (qs when L > 0)
extern "C" int dn_read_model(const char *mfile, int nx, int ny, int L, float *vs, float *rho, float *qs)
{
    const std::string base = mfile ? mfile : "";
    if (dn_read_model_file((base + ".vs").c_str(), nx, ny, vs)) return -1;
    if (dn_read_model_file((base + ".rho").c_str(), nx, ny, rho)) return -1;
    if (L > 0 && qs && dn_read_model_file((base + ".qs").c_str(), nx, ny, qs)) return -1;
    return 0;
}

namespace {
template <typename T> void put(unsigned char *h, int off, T v) { std::memcpy(h + off, &v, sizeof(T)); }
}  // namespace

// outseis_glob.c:46-265 (SEIS_FORMAT 1): trace header fields and samples.
extern "C" int dn_write_su(const char *path, int ntr, int ns, const float *data, const int *rec_i, const int *rec_j,
                           float dh, float refrec_x, float refrec_y, float src_x, float src_y, float dt, int ndt)
{
    FILE *fp = std::fopen(path ? path : "", "wb");
    if (!fp) return dn_fail(nullptr, "Could not open %s for writing", path ? path : "(null)");
    const double kPi = 3.141592653589793;
    const float xshift = 800.0f, yshift = 800.0f;
    const int step = ndt < 1 ? 1 : ndt;
    const int ns2 = ns / step;  // outseis_glob.c:32 (INVMAT 10): every NDT-th sample
    for (int t = 0; t < ntr; ++t) {
        unsigned char h[240];
        std::memset(h, 0, sizeof(h));
        const float xr = rec_i[t] * dh, yr = rec_j[t] * dh;
        const float x = xr - refrec_x, y = yr - refrec_y;
        // tracl and cdp carry recpos[3] = iround(REFREC[3]/DH) = 0 in the reference
        put<short>(h, 28, 1);                                                              // trid
        put<int>(h, 36, iroundd(std::sqrt((double)(((src_x - xr) * (src_x - xr)) + ((src_y - yr) * (src_y - yr)))) * 1000.0));
        put<int>(h, 40, iroundd((double)yr * 1000.0));                                    // gelev
        put<int>(h, 48, iroundd((double)src_y * 1000.0));                                 // sdepth
        put<int>(h, 60, iroundd(((360.0 / (2.0 * kPi)) * std::atan2((double)(x - xshift), (double)(y - yshift))) * 1000.0));
        put<short>(h, 68, (short)-1000);                                                   // scalel
        put<short>(h, 70, (short)-1000);                                                   // scalco
        put<int>(h, 72, iroundd((double)src_x * 1000.0));                                 // sx
        put<int>(h, 76, iroundd((double)src_y * 1000.0));                                 // sy
        put<int>(h, 80, iroundd((double)xr * 1000.0));                                    // gx
        put<int>(h, 84, iroundd((double)yr * 1000.0));                                    // gy
        put<unsigned short>(h, 114, (unsigned short)ns2);                                  // ns
        const unsigned short dtus = (unsigned short)iroundd((double)((float)step * dt) * 1.0e6);
        put<unsigned short>(h, 116, dtus);                                                 // dt (us)
        put<float>(h, 180, (float)((double)(float)dtus * 1.0e-6));                         // d1
        std::fwrite(h, 1, 240, fp);
        for (int k = 1; k <= ns2; ++k) {
            const float v = data[(size_t)t * ns + (size_t)(k * step - 1)];
            std::fwrite(&v, sizeof(float), 1, fp);
        }
    }
    std::fclose(fp);
    return 0;
}

// inseis.c:71-84: out[t][0] = 0 and out[t][k] = file sample k (0-based,
// k >= 1): the reference zeroes the first observed sample of every trace.
extern "C" int dn_read_su(const char *path, int ntr, int ns, float *out)
{
    FILE *fp = std::fopen(path ? path : "", "rb");
    if (!fp) return dn_fail(nullptr, "Field data file could not be opened: %s", path ? path : "(null)");
    std::vector<float> tr((size_t)ns);
    for (int t = 0; t < ntr; ++t) {
        unsigned char h[240];
        if (std::fread(h, 1, 240, fp) != 240 || std::fread(tr.data(), sizeof(float), (size_t)ns, fp) != (size_t)ns) {
            std::fclose(fp);
            return dn_fail(nullptr, "SU file %s: fewer than %d traces of %d samples", path, ntr, ns);
        }
        float *row = out + (size_t)t * ns;
        row[0] = 0.0f;
        for (int k = 1; k < ns; ++k) row[k] = tr[k];
    }
    std::fclose(fp);
    return 0;
}
