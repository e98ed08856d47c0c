// dq_io.cpp -- the files of the reference's on-disk workflow: Hamiltonian,
// initial state, dipoles, per-site noise parameters, spectral density,
// infile.in, and the population file with its continuation mode.  Host code
// only; every reader validates what it reads and reports through dq_status.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/dynaquant.h"

namespace dq {
dq_status set_error(dq_status st, const std::string& msg);   // dq_api.cpp: the thread's dq_last_error
}

namespace {

dq_status fail(dq_status st, const std::string& msg) { return dq::set_error(st, msg); }

std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

bool read_base_line(std::ifstream& f, int* base)
{
    std::string line;
    if (!std::getline(f, line)) return false;
    line = trim(line);
    if (line == "BASE_ZERO") { *base = 0; return true; }
    if (line == "BASE_ONE") { *base = 1; return true; }
    return false;
}

// rows of whitespace-separated numbers; blank lines are skipped
bool read_numeric_rows(const char* path, int ncol, std::vector<double>* out, std::string* err)
{
    std::ifstream f(path);
    if (!f.is_open()) { *err = std::string("cannot open ") + path; return false; }
    std::string line;
    int lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        if (trim(line).empty()) continue;
        std::istringstream ss(line);
        for (int c = 0; c < ncol; ++c) {
            double v;
            if (!(ss >> v) || !std::isfinite(v)) {
                *err = std::string(path) + ": line " + std::to_string(lineno) + " needs " + std::to_string(ncol) + " numbers";
                return false;
            }
            out->push_back(v);
        }
    }
    return true;
}

void copy_str(char* dst, const std::string& s)
{
    std::strncpy(dst, s.c_str(), 511);
    dst[511] = 0;
}

}  // namespace

extern "C" {

// sparseMatrixClasses.cu:74-267 (format), :489-540 (COO -> CSR)
dq_status dq_read_hamiltonian_file(const char* path, int32_t* n_out, int64_t* nnz_out, int32_t* rowptr,
                                   int32_t* col, double* val, double* e_site)
{
    if (!path || !n_out || !nnz_out) return fail(DQ_ERR_INVALID_ARG, "null argument");
    std::ifstream f(path);
    if (!f.is_open()) return fail(DQ_ERR_INVALID_ARG, std::string("cannot open ") + path);
    std::string line, a, b, c, field, symm;
    if (!std::getline(f, line)) return fail(DQ_ERR_INVALID_ARG, "empty Hamiltonian file");
    {
        std::istringstream ss(line);
        ss >> a >> b >> c >> field >> symm;
        for (auto* s : {&field, &symm}) std::transform(s->begin(), s->end(), s->begin(), ::tolower);
        if (field != "real" && field != "complex")
            return fail(DQ_ERR_INVALID_ARG, "Hamiltonian file: value type must be real");
        if (symm != "general" && symm != "symmetric" && symm != "hermitian")
            return fail(DQ_ERR_INVALID_ARG, "Hamiltonian file: matrix type must be general, symmetric or hermitian");
    }
    int base = 0;
    if (!read_base_line(f, &base)) return fail(DQ_ERR_INVALID_ARG, "Hamiltonian file: second line must be BASE_ZERO or BASE_ONE");
    int64_t n = 0, n2 = 0, nnz = 0;
    if (!std::getline(f, line)) return fail(DQ_ERR_INVALID_ARG, "Hamiltonian file: missing size line");
    {
        std::istringstream ss(line);
        if (!(ss >> n >> n2 >> nnz) || n <= 0 || n2 != n || nnz <= 0 || n > (1 << 30))
            return fail(DQ_ERR_INVALID_ARG, "Hamiltonian file: size line must be 'N N nnz' with N, nnz > 0");
    }
    const bool mirror = symm != "general";
    std::vector<double> diag(n, 0.0);
    std::vector<std::map<int32_t, double>> rows(n);
    int64_t count = 0;
    while (std::getline(f, line)) {
        if (trim(line).empty()) continue;
        std::istringstream ss(line);
        int64_t i, j;
        double v;
        if (!(ss >> i >> j >> v) || !std::isfinite(v))
            return fail(DQ_ERR_INVALID_ARG, "Hamiltonian file: entry lines must be 'row col value'");
        i -= base;
        j -= base;
        if (i < 0 || i >= n || j < 0 || j >= n) return fail(DQ_ERR_INVALID_ARG, "Hamiltonian file: index out of range");
        if (++count > nnz) return fail(DQ_ERR_INVALID_ARG, "Hamiltonian file: more entries than nnz");
        if (i == j) { diag[i] += v; continue; }
        rows[i][(int32_t)j] += v;
        if (mirror) rows[j][(int32_t)i] += v;
    }
    if (count != nnz) return fail(DQ_ERR_INVALID_ARG, "Hamiltonian file: fewer entries than nnz");
    if (!mirror) {
        for (int64_t i = 0; i < n; ++i)
            for (const auto& kv : rows[i]) {
                auto it = rows[kv.first].find((int32_t)i);
                if (it == rows[kv.first].end() || it->second != kv.second)
                    return fail(DQ_ERR_INVALID_ARG, "Hamiltonian file: a general matrix must be symmetric");
            }
    }
    int64_t off = 0;
    for (const auto& r : rows) off += (int64_t)r.size();
    *n_out = (int32_t)n;
    *nnz_out = off;
    if (!rowptr) return DQ_OK;
    if (!col || !val || !e_site) return fail(DQ_ERR_INVALID_ARG, "null CSR output");
    int64_t p = 0;
    rowptr[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        for (const auto& kv : rows[i]) { col[p] = kv.first; val[p] = kv.second; ++p; }
        rowptr[i + 1] = (int32_t)p;
        e_site[i] = diag[i];
    }
    return DQ_OK;
}

// systemClass.cu:1383-1481 (coefficients read as double, not float; reading
// stops at END or at the end of the file)
dq_status dq_read_state_file(const char* path, int32_t n, double* psi)
{
    if (!path || !psi || n <= 0) return fail(DQ_ERR_INVALID_ARG, "null argument or n <= 0");
    std::ifstream f(path);
    if (!f.is_open()) return fail(DQ_ERR_INVALID_ARG, std::string("cannot open ") + path);
    int base = 0;
    if (!read_base_line(f, &base)) return fail(DQ_ERR_INVALID_ARG, "state file: first line must be BASE_ZERO or BASE_ONE");
    std::string line;
    std::getline(f, line);                      // comment line
    std::vector<double> re(n, 0.0);
    while (std::getline(f, line)) {
        if (line.find("END") != std::string::npos) break;   // END is optional, as in the reference
        if (trim(line).empty()) continue;
        std::istringstream ss(line);
        int64_t i;
        double v;
        if (!(ss >> i >> v) || !std::isfinite(v)) return fail(DQ_ERR_INVALID_ARG, "state file: lines must be 'index coefficient'");
        i -= base;
        if (i < 0 || i >= n) return fail(DQ_ERR_INVALID_ARG, "state file: index out of range");
        re[i] = v;
    }
    double s = 0.0;
    for (double v : re) s += v * v;
    if (!(s > 0.0)) return fail(DQ_ERR_INVALID_ARG, "state file: zero state");
    const double inv = 1.0 / std::sqrt(s);
    for (int i = 0; i < n; ++i) { psi[2 * i] = re[i] * inv; psi[2 * i + 1] = 0.0; }
    return DQ_OK;
}

// spectraClass.cu:173-262
dq_status dq_read_dipole_file(const char* path, int32_t n, double* mu)
{
    if (!path || !mu || n <= 0) return fail(DQ_ERR_INVALID_ARG, "null argument or n <= 0");
    std::vector<double> v;
    std::string err;
    if (!read_numeric_rows(path, 3, &v, &err)) return fail(DQ_ERR_INVALID_ARG, err);
    if ((int64_t)v.size() != 3 * (int64_t)n) return fail(DQ_ERR_INVALID_ARG, "dipole file: needs exactly n lines");
    std::copy(v.begin(), v.end(), mu);
    return DQ_OK;
}

// posfile, read line by line beside the dipole file (spectraClass.cu:296-360)
dq_status dq_read_position_file(const char* path, int32_t n, double* pos)
{
    if (!path || !pos || n <= 0) return fail(DQ_ERR_INVALID_ARG, "null argument or n <= 0");
    std::vector<double> v;
    std::string err;
    if (!read_numeric_rows(path, 3, &v, &err)) return fail(DQ_ERR_INVALID_ARG, err);
    if ((int64_t)v.size() != 3 * (int64_t)n) return fail(DQ_ERR_INVALID_ARG, "position file: needs exactly n lines");
    std::copy(v.begin(), v.end(), pos);
    return DQ_OK;
}

// R x mu per site and the norms of each Cartesian component over the sites
// (spectraClass.cu:346-411).  The reference adds the imaginary part of its
// complex dipoles into the cross product; its dipoles are real, so that term
// is zero and is left out here.
dq_status dq_cd_vectors(int32_t n, const double* pos, const double* mu, double* rd, double* mu_tot,
                        double* rd_tot)
{
    if (!pos || !mu || !rd || !mu_tot || !rd_tot || n <= 0) return fail(DQ_ERR_INVALID_ARG, "null argument or n <= 0");
    double m2[3] = {0.0, 0.0, 0.0}, r2[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < n; ++i) {
        const double* r = pos + 3 * (size_t)i;
        const double* u = mu + 3 * (size_t)i;
        double* c = rd + 3 * (size_t)i;
        c[0] = r[1] * u[2] - r[2] * u[1];
        c[1] = r[2] * u[0] - r[0] * u[2];
        c[2] = r[0] * u[1] - r[1] * u[0];
        for (int d = 0; d < 3; ++d) {
            m2[d] += u[d] * u[d];
            r2[d] += c[d] * c[d];
        }
    }
    for (int d = 0; d < 3; ++d) {
        mu_tot[d] = std::sqrt(m2[d]);
        rd_tot[d] = std::sqrt(r2[d]);
    }
    return DQ_OK;
}

// ka_paramsfile "tau sigma" (systemClass.cu:980-1071), hsr_stddevfile "sigma"
// (systemClass.cu:910-975)
dq_status dq_read_site_columns(const char* path, int32_t n, int32_t ncol, double* out)
{
    if (!path || !out || n <= 0 || ncol < 1 || ncol > 8) return fail(DQ_ERR_INVALID_ARG, "bad argument");
    std::vector<double> v;
    std::string err;
    if (!read_numeric_rows(path, ncol, &v, &err)) return fail(DQ_ERR_INVALID_ARG, err);
    if ((int64_t)v.size() != (int64_t)ncol * n) return fail(DQ_ERR_INVALID_ARG, "site parameter file: needs exactly n lines");
    std::copy(v.begin(), v.end(), out);
    return DQ_OK;
}

// systemClass.cu:1090-1160: first line "domega nosc", then "omega J" lines
dq_status dq_read_spectral_density(const char* path, int32_t* nosc_out, double* domega_out, double* omega, double* J)
{
    if (!path || !nosc_out || !domega_out) return fail(DQ_ERR_INVALID_ARG, "null argument");
    std::ifstream f(path);
    if (!f.is_open()) return fail(DQ_ERR_INVALID_ARG, std::string("cannot open ") + path);
    std::string line;
    double dw = 0.0;
    int64_t nosc = 0;
    if (!std::getline(f, line)) return fail(DQ_ERR_INVALID_ARG, "empty spectral density file");
    {
        std::istringstream ss(line);
        if (!(ss >> dw >> nosc) || !(dw > 0.0) || nosc <= 0 || nosc > (1 << 20))
            return fail(DQ_ERR_INVALID_ARG, "spectral density file: first line must be 'domega nosc'");
    }
    std::vector<double> w, j;
    while (std::getline(f, line)) {
        if (trim(line).empty()) continue;
        std::istringstream ss(line);
        double a, b;
        if (!(ss >> a >> b) || !std::isfinite(a) || !std::isfinite(b))
            return fail(DQ_ERR_INVALID_ARG, "spectral density file: lines must be 'omega J'");
        w.push_back(a);
        j.push_back(b);
    }
    if ((int64_t)w.size() != nosc) return fail(DQ_ERR_INVALID_ARG, "spectral density file: line count differs from nosc");
    *nosc_out = (int32_t)nosc;
    *domega_out = dw;
    if (omega) std::copy(w.begin(), w.end(), omega);
    if (J) std::copy(j.begin(), j.end(), J);
    return DQ_OK;
}

// systemClass.cu:183-735.  "key = value" lines; a key containing '!' is a
// comment; numbers are read with atof/atoi, so text after them is ignored.
dq_status dq_parse_infile(const char* path, dq_infile* out)
{
    if (!path || !out) return fail(DQ_ERR_INVALID_ARG, "null argument");
    std::ifstream f(path);
    if (!f.is_open()) return fail(DQ_ERR_INVALID_ARG, std::string("cannot open ") + path);
    dq_infile r;
    std::memset(&r, 0, sizeof(r));
    r.dt = 1.0;
    r.m = 12;
    r.ka_substeps = 1;
    r.tk_p = 10;                // systemClass.cu:70-71
    r.tk_m_aux = 24;
    r.steps_per_loop = 0;
    r.noise_model = DQ_NOISE_NONE;
    r.gauss_width = 1.0;
    for (int k = 0; k < 10; ++k) r.print_pops[k] = -1;
    std::string line;
    int lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        const std::string key = trim(line.substr(0, eq));
        if (key.find('!') != std::string::npos) continue;
        std::string v = trim(line.substr(eq + 1));
        auto flag = [&](int32_t* dst) -> bool {
            const int x = std::atoi(v.c_str());
            if (x != 0 && x != 1) return false;
            *dst = x;
            return true;
        };
        auto word = [&]() {
            // string values end at the first blank or '!'
            const size_t e = v.find_first_of(" \t!");
            return e == std::string::npos ? v : v.substr(0, e);
        };
        bool ok = true;
        if (key == "dt") r.dt = std::atof(v.c_str());
        else if (key == "totalsteps") r.total_steps = std::atoi(v.c_str());
        else if (key == "ensembleruns") r.ensemble_runs = std::atoi(v.c_str());
        else if (key == "m") r.m = std::atoi(v.c_str());
        else if (key == "ka_substeps") r.ka_substeps = std::atoi(v.c_str());
        else if (key == "temperature") r.temperature = std::atof(v.c_str());
        else if (key == "sigmastatdis") r.sigma_static = std::atof(v.c_str());
        else if (key == "gausswidth") r.gauss_width = std::atof(v.c_str());
        else if (key == "stepsperloop") r.steps_per_loop = std::atoi(v.c_str());
        else if (key == "initstate") r.init_state = std::atoi(v.c_str());
        else if (key == "debug_diags") ok = flag(&r.debug_diags);
        else if (key == "print_partic") ok = flag(&r.print_partic);
        else if (key == "dostatelooping") ok = flag(&r.state_looping);
        else if (key == "writeoutpops") ok = flag(&r.write_pops);
        else if (key == "writebinarypops") ok = flag(&r.write_binary_pops);
        else if (key == "printallpops") ok = flag(&r.print_all_pops);
        else if (key == "writeoutstates") r.write_states = std::atoi(v.c_str());
        else if (key == "staticdis") ok = flag(&r.static_dis);
        else if (key == "absorbspec") ok = flag(&r.absorb_spec);
        else if (key == "cdspec") ok = flag(&r.cd_spec);
        else if (key == "lineshapedress") ok = flag(&r.lineshape_dress);
        else if (key == "initfromfile") ok = flag(&r.init_from_file);
        else if (key.size() == 9 && key.compare(0, 8, "printpop") == 0 && std::isdigit((unsigned char)key[8]))
            r.print_pops[key[8] - '0'] = std::atoi(v.c_str());
        else if (key == "ensembleavgmode") {
            const std::string w = word();
            r.ensemble_mode = w == "ENS_AVG_MODE_FIRST" ? 1 : (w == "ENS_AVG_MODE_CONTINUATION" ? 2 : 0);
        } else if (key == "simtype") {
            const std::string w = word();
            if (w == "SIM_TYPE_CONST_HAM") r.noise_model = DQ_NOISE_NONE;
            else if (w == "SIM_TYPE_KUBO_ANDERSON") r.noise_model = DQ_NOISE_OU;
            else if (w == "SIM_TYPE_HSR") r.noise_model = DQ_NOISE_HSR;
            else if (w == "SIM_TYPE_ZZ_REAL_NOISE") r.noise_model = DQ_NOISE_ZZ;
            else ok = false;
        } else if (key == "propmode") {
            const std::string w = word();
            if (w == "PROP_MODE_LANCZOS") r.prop_mode = 0;
            else if (w == "PROP_MODE_TAYLOR_LANCZOS") r.prop_mode = 1;
            else ok = false;
        } else if (key == "hamfile") copy_str(r.ham_file, word());
        else if (key == "dipolefile") copy_str(r.dipole_file, word());
        else if (key == "posfile") copy_str(r.pos_file, word());
        else if (key == "tkmatfile") copy_str(r.tk_mat_file, word());
        else if (key == "tk_p") r.tk_p = std::atoi(v.c_str());
        else if (key == "tk_m_aux") r.tk_m_aux = std::atoi(v.c_str());
        else if (key == "initstatefile") copy_str(r.init_state_file, word());
        else if (key == "ka_paramsfile") copy_str(r.ka_params_file, word());
        else if (key == "hsr_stddevfile") copy_str(r.hsr_file, word());
        else if (key == "zz_specd_file") copy_str(r.zz_specd_file, word());
        else if (key == "outfolder") copy_str(r.out_folder, word());
        if (!ok) return fail(DQ_ERR_INVALID_ARG, std::string(path) + ": line " + std::to_string(lineno) + ": bad value for " + key);
    }
    if (r.absorb_spec && r.cd_spec) return fail(DQ_ERR_INVALID_ARG, "absorbspec and cdspec are exclusive");
    *out = r;
    return DQ_OK;
}

// popPrintClass.cu:159-335 (headers, continuation), :442-472 (values), :490-522
// (temporary file renamed over the old one).  pops[t][site] for steps 1..T.
dq_status dq_read_pops_file(const char* path, int32_t binary, int32_t* nsims, double* dt, int32_t* total_steps,
                            int32_t* nsites, double* pops)
{
    if (!path || !nsims || !dt || !total_steps || !nsites) return fail(DQ_ERR_INVALID_ARG, "null argument");
    if (binary) {
        FILE* f = std::fopen(path, "rb");
        if (!f) return fail(DQ_ERR_INVALID_ARG, std::string("cannot open ") + path);
        int32_t a, c, e;
        double b;
        bool ok = std::fread(&a, 4, 1, f) == 1 && std::fread(&b, 8, 1, f) == 1 && std::fread(&c, 4, 1, f) == 1 &&
                  std::fread(&e, 4, 1, f) == 1 && a > 0 && c >= 0 && e >= 0;
        if (ok && pops) ok = std::fread(pops, 8, (size_t)c * e, f) == (size_t)c * e;
        std::fclose(f);
        if (!ok) return fail(DQ_ERR_INVALID_ARG, "population file: truncated or bad header");
        *nsims = a; *dt = b; *total_steps = c; *nsites = e;
        return DQ_OK;
    }
    std::ifstream f(path);
    if (!f.is_open()) return fail(DQ_ERR_INVALID_ARG, std::string("cannot open ") + path);
    const char* keys[4] = {"numsims", "dt", "totalsteps", "numprintedsites"};
    double hv[4];
    std::string line;
    for (int k = 0; k < 4; ++k) {
        std::string key;
        if (!std::getline(f, line)) return fail(DQ_ERR_INVALID_ARG, "population file: truncated header");
        std::istringstream ss(line);
        if (!(ss >> key >> hv[k]) || key != keys[k])
            return fail(DQ_ERR_INVALID_ARG, std::string("population file: expected ") + keys[k]);
    }
    *nsims = (int32_t)hv[0]; *dt = hv[1]; *total_steps = (int32_t)hv[2]; *nsites = (int32_t)hv[3];
    if (!pops) return DQ_OK;
    for (int32_t t = 0; t < *total_steps; ++t) {
        if (!std::getline(f, line) || line.compare(0, 4, "STEP") != 0)
            return fail(DQ_ERR_INVALID_ARG, "population file: missing STEP line");
        for (int32_t s = 0; s < *nsites; ++s) {
            if (!std::getline(f, line)) return fail(DQ_ERR_INVALID_ARG, "population file: truncated");
            std::istringstream ss(line);
            if (!(ss >> pops[(size_t)t * *nsites + s])) return fail(DQ_ERR_INVALID_ARG, "population file: bad value");
        }
    }
    return DQ_OK;
}

dq_status dq_write_pops_file(const char* path, int32_t binary, double dt, int32_t total_steps, int32_t nsites,
                             const double* pops, int32_t nsims)
{
    if (!path || !pops || total_steps < 1 || nsites < 1 || nsims < 1) return fail(DQ_ERR_INVALID_ARG, "bad argument");
    const size_t count = (size_t)total_steps * nsites;
    std::vector<double> sum(pops, pops + count);
    int32_t total_sims = nsims;
    // the header's dt as the text file stores it (6 significant digits)
    char dts[64];
    std::snprintf(dts, sizeof(dts), "%g", dt);
    const double dt_text = std::atof(dts);
    if (FILE* probe = std::fopen(path, "rb")) {
        std::fclose(probe);
        int32_t pn, pt, ps;
        double pdt;
        dq_status st = dq_read_pops_file(path, binary, &pn, &pdt, &pt, &ps, nullptr);
        if (st != DQ_OK) return st;
        if (pt != total_steps || ps != nsites || pdt != (binary ? dt : dt_text))
            return fail(DQ_ERR_INVALID_ARG, "population file: dt, totalsteps or numprintedsites differ from the existing file");
        std::vector<double> prev(count);
        if ((st = dq_read_pops_file(path, binary, &pn, &pdt, &pt, &ps, prev.data())) != DQ_OK) return st;
        for (size_t k = 0; k < count; ++k) sum[k] += prev[k];
        total_sims += pn;
    }
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), binary ? "wb" : "w");
    if (!f) return fail(DQ_ERR_INVALID_ARG, "cannot create " + tmp);
    bool ok = true;
    if (binary) {
        ok = std::fwrite(&total_sims, 4, 1, f) == 1 && std::fwrite(&dt, 8, 1, f) == 1 &&
             std::fwrite(&total_steps, 4, 1, f) == 1 && std::fwrite(&nsites, 4, 1, f) == 1 &&
             std::fwrite(sum.data(), 8, count, f) == count;
    } else {
        ok = std::fprintf(f, "numsims %d\ndt %s\ntotalsteps %d\nnumprintedsites %d\n", total_sims, dts, total_steps,
                          nsites) > 0;
        for (int32_t t = 0; ok && t < total_steps; ++t) {
            ok = std::fprintf(f, "STEP %d\n", t + 1) > 0;
            for (int32_t s = 0; ok && s < nsites; ++s)
                ok = std::fprintf(f, "%25.17g (%d)\n", sum[(size_t)t * nsites + s], s) > 0;
        }
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return fail(DQ_ERR_INVALID_ARG, std::string("writing ") + path + " failed");
    }
    return DQ_OK;
}

}  // extern "C"
