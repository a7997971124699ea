// Output files of the reference driver, in its format: one ".dat" file per
// series under the output directory, every number in scientific notation with
// 15 digits in a column of width 24 (src/Utils.h:25-28, src/Utils.cpp:788-816),
// so analysis scripts written for the reference read these files unchanged.
//   - vector files: a header line, then one value per line
//     (WriteDataToFile, src/Utils.cpp:829-841);
//   - series files: one line of values per call, appended
//     (AppendDataToFile, src/Utils.cpp:1000-1016).
// Host code only; nothing here touches the device.

#include "../../include/tdvmc.h"
#include "config.hpp"

#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace {

// "%24.15e" is what an ostream in scientific mode with precision 15 and
// setw(24) prints for a double
int put_values(std::FILE* f, const double* v, int n, bool one_per_line)
{
    for (int i = 0; i < n; i++) {
        if (std::fprintf(f, "%24.15e", v[i]) < 0) return -1;
        if (one_per_line && std::fputc('\n', f) == EOF) return -1;
    }
    if (!one_per_line && std::fputc('\n', f) == EOF) return -1;
    return 0;
}

std::string path_of(const char* dir, const char* name) { return std::string(dir ? dir : "") + name + ".dat"; }

// WriteDataToFile(vector, name, header): truncate, header, one value per line
int write_vector(const char* dir, const char* name, const char* header, const double* v, int n)
{
    std::FILE* f = std::fopen(path_of(dir, name).c_str(), "w");
    if (!f) return TDVMC_ERR_BAD_ARG;
    int rc = std::fprintf(f, "%s\n", header) < 0 ? -1 : put_values(f, v, n, true);
    rc |= std::fclose(f);
    return rc ? TDVMC_ERR_BAD_ARG : TDVMC_OK;
}

// AppendDataToFile(vector or scalar, name): one line at the end of the file
int append_line(const char* dir, const char* name, const double* v, int n)
{
    std::FILE* f = std::fopen(path_of(dir, name).c_str(), "a");
    if (!f) return TDVMC_ERR_BAD_ARG;
    int rc = put_values(f, v, n, false);
    rc |= std::fclose(f);
    return rc ? TDVMC_ERR_BAD_ARG : TDVMC_OK;
}

// GetCoordinateNIC (src/Utils.cpp:266-281): x - k L with k the quotient
// x (1/L) rounded half away from zero, and a result of exactly +-L/2 nudged
// inwards by 1e-10 (the same arithmetic as there: 1/L and L/2 precomputed)
double nic_coordinate(double x, double L)
{
    const double Lr = 1.0 / L, L2 = L / 2.0;
    const int k = (int)(x * Lr + ((x >= 0.0) ? 0.5 : -0.5));
    double r = x - k * L;
    if (r == L2)
        r -= 1e-10;
    else if (r == -L2)
        r += 1e-10;
    return r;
}

int ensure_dir(const std::string& d)
{
    struct stat st;
    if (stat(d.c_str(), &st) == 0) return S_ISDIR(st.st_mode) ? 0 : -1;
    return mkdir(d.c_str(), 0755);
}

}  // namespace

extern "C" {

int tdvmc_files_write_coordinates(const char* out_dir, const double* R, int64_t W, int N, int D, double L)
{
    if (!out_dir || !R || W < 1 || N < 1 || D < 1 || D > 3) return TDVMC_ERR_BAD_ARG;
    const std::string dir(out_dir);
    if (ensure_dir(dir + "coords")) return TDVMC_ERR_BAD_ARG;
    std::vector<double> row((size_t)N * D);
    for (int64_t w = 0; w < W; w++) {
        // AlignCoordinates (src/TDVMC.cpp:2569-2582, :787-796): into the first cell
        for (size_t i = 0; i < row.size(); i++) {
            const double x = R[(size_t)w * row.size() + i];
            row[i] = L > 0 ? nic_coordinate(x, L) : x;
        }
        const std::string name = "coords/particleconfiguration_" + std::to_string(N) + "_" + std::to_string(D) +
                                 "D_" + std::to_string((long long)w);
        std::FILE* f = std::fopen(path_of(out_dir, name.c_str()).c_str(), "w");
        if (!f) return TDVMC_ERR_BAD_ARG;
        int rc = put_values(f, row.data(), (int)row.size(), false);
        rc |= std::fclose(f);
        if (rc) return TDVMC_ERR_BAD_ARG;
        if (w == 0) {
            // the root rank's copy for the next run's config directory (:4019-4022)
            const std::string fin = "AAFinish_particleconfiguration_" + std::to_string(N) + "_" + std::to_string(D) + "D";
            std::FILE* g = std::fopen(path_of(out_dir, fin.c_str()).c_str(), "w");
            if (!g) return TDVMC_ERR_BAD_ARG;
            int rc2 = put_values(g, row.data(), (int)row.size(), false);
            rc2 |= std::fclose(g);
            if (rc2) return TDVMC_ERR_BAD_ARG;
        }
    }
    return TDVMC_OK;
}

int tdvmc_files_read_coordinates(const char* path, double* R, int N, int D)
{
    if (!path || !R || N < 1 || D < 1) return TDVMC_ERR_BAD_ARG;
    std::FILE* f = std::fopen(path, "r");
    if (!f) return TDVMC_ERR_BAD_ARG;
    // the last line of the file (LoadLastPositionsFromFile, src/TDVMC.cpp:640-670)
    std::string line, last;
    int c;
    while ((c = std::fgetc(f)) != EOF) {
        if (c == '\n') {
            if (!line.empty()) last = line;
            line.clear();
        } else if (c != '\r') {
            line.push_back((char)c);
        }
    }
    if (!line.empty()) last = line;
    std::fclose(f);
    // comma-separated values, DIM per particle in particle order
    std::vector<double> v;
    const char* p = last.c_str();
    while (*p) {
        char* end = nullptr;
        const double x = std::strtod(p, &end);
        if (end == p) break;
        v.push_back(x);
        p = end;
        while (*p == ',' || *p == ' ' || *p == '\t') p++;
    }
    if ((int64_t)v.size() != (int64_t)N * D) return TDVMC_ERR_BAD_ARG;
    for (size_t i = 0; i < v.size(); i++) R[i] = v[i];
    return TDVMC_OK;
}

int tdvmc_files_write_rng_state(const char* out_dir, uint64_t seed, uint32_t epoch, int rank)
{
    if (!out_dir || rank < 0) return TDVMC_ERR_BAD_ARG;
    const std::string dir(out_dir);
    if (ensure_dir(dir + "random")) return TDVMC_ERR_BAD_ARG;
    const std::string path = dir + "random/state_philox_" + std::to_string(rank) + ".dat";
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return TDVMC_ERR_BAD_ARG;
    int rc = std::fprintf(f, "philox4x32-10 %" PRIu64 " %" PRIu32 "\n", seed, epoch) < 0 ? -1 : 0;
    rc |= std::fclose(f);
    return rc ? TDVMC_ERR_BAD_ARG : TDVMC_OK;
}

int tdvmc_files_read_rng_state(const char* path, uint64_t* seed, uint32_t* epoch)
{
    if (!path || !seed || !epoch) return TDVMC_ERR_BAD_ARG;
    std::FILE* f = std::fopen(path, "r");
    if (!f) return TDVMC_ERR_BAD_ARG;
    char name[32] = {0};
    uint64_t s = 0;
    uint32_t e = 0;
    const int n = std::fscanf(f, "%31s %" SCNu64 " %" SCNu32, name, &s, &e);
    std::fclose(f);
    if (n != 3 || std::string(name) != "philox4x32-10") return TDVMC_ERR_BAD_ARG;
    *seed = s;
    *epoch = e;
    return TDVMC_OK;
}

int tdvmc_files_init(const char* out_dir, const double* uR, const double* uI, int P)
{
    if (!out_dir || P < 0 || (P > 0 && (!uR || !uI))) return TDVMC_ERR_BAD_ARG;
    // src/TDVMC.cpp:3118-3129: start parameters, then the series files with their headers only
    int rc = write_vector(out_dir, "ParametersR0", "uR_k...", uR, P);
    if (!rc) rc = write_vector(out_dir, "ParametersI0", "uI_k...", uI, P);
    if (!rc) rc = write_vector(out_dir, "LocalEnergyR", "E^R", nullptr, 0);
    if (!rc) rc = write_vector(out_dir, "LocalEnergyI", "E^I", nullptr, 0);
    if (!rc) rc = write_vector(out_dir, "LocalOperators", "<O_k>...", nullptr, 0);
    if (!rc) rc = write_vector(out_dir, "OtherExpectationValues", "kinetic, potential, wf, g(r)", nullptr, 0);
    if (!rc) rc = write_vector(out_dir, "ParametersR", "uR_k...", nullptr, 0);
    if (!rc) rc = write_vector(out_dir, "ParametersI", "uI_k...", nullptr, 0);
    if (!rc) rc = write_vector(out_dir, "times", "t", nullptr, 0);
    return rc;
}

int tdvmc_files_append_step(const char* out_dir, double time, const tdvmc_step_data* d, const double* uR,
                            const double* uI, int P)
{
    if (!out_dir || !d || P < 0 || (P > 0 && (!uR || !uI || !d->localOperators)) || !d->otherExpectationValues)
        return TDVMC_ERR_BAD_ARG;
    // src/TDVMC.cpp:3833-3839
    int rc = append_line(out_dir, "timesSystem", &time, 1);
    if (!rc) rc = append_line(out_dir, "LocalEnergyR", &d->localEnergyR, 1);
    if (!rc) rc = append_line(out_dir, "LocalEnergyI", &d->localEnergyI, 1);
    if (!rc) rc = append_line(out_dir, "LocalOperators", d->localOperators, P);
    if (!rc) rc = append_line(out_dir, "OtherExpectationValues", d->otherExpectationValues, 9);
    if (!rc) rc = append_line(out_dir, "ParametersR", uR, P);
    if (!rc) rc = append_line(out_dir, "ParametersI", uI, P);
    return rc;
}

int tdvmc_files_write_grids(const char* out_dir, const double* gr_grid, int n_gr, const double* sk_grid, int n_sk)
{
    if (!out_dir || n_gr < 0 || n_sk < 0 || (n_gr > 0 && !gr_grid) || (n_sk > 0 && !sk_grid)) return TDVMC_ERR_BAD_ARG;
    // src/TDVMC.cpp:3151-3157: the grids' names are the headers (BosonsBulk.cpp:143, :150)
    int rc = write_vector(out_dir, "gr_grid", "r_ij", gr_grid, n_gr);
    if (!rc) rc = write_vector(out_dir, "sk_grid", "k", sk_grid, n_sk);
    return rc;
}

int tdvmc_files_append_observables(const char* out_dir, double time, const double* gr, int n_gr, const double* sk,
                                   int n_sk)
{
    if (!out_dir || n_gr < 0 || n_sk < 0 || (n_gr > 0 && !gr) || (n_sk > 0 && !sk)) return TDVMC_ERR_BAD_ARG;
    // src/TDVMC.cpp:3475, :3496-3500
    int rc = append_line(out_dir, "timesAdditional", &time, 1);
    if (!rc) rc = append_line(out_dir, "gr", gr, n_gr);
    if (!rc) rc = append_line(out_dir, "sk", sk, n_sk);
    return rc;
}

int tdvmc_write_final_config(const char* config, const double* uR, const double* uI, double phiR, double phiI, int P,
                             const char* path)
{
    if (!config || !path || P < 0 || (P > 0 && (!uR || !uI))) return TDVMC_ERR_BAD_ARG;
    std::string text;
    try {
        // src/TDVMC.cpp:4026-4034: the final parameters go into PARAMS_REAL,
        // PARAMS_IMAGINARY, PARAM_PHIR and PARAM_PHII
        tdvmc::Config c = tdvmc::parse_config(tdvmc::read_config_text(config));
        c.PARAMS_REAL.assign(uR, uR + P);
        c.PARAMS_IMAGINARY.assign(uI, uI + P);
        c.PARAM_PHIR = phiR;
        c.PARAM_PHII = phiI;
        text = tdvmc::config_to_json(c);
    } catch (const std::exception&) {
        return TDVMC_ERR_BAD_CONFIG;
    }
    std::FILE* f = std::fopen(path, "w");
    if (!f) return TDVMC_ERR_BAD_ARG;
    int rc = std::fputs(text.c_str(), f) < 0 ? -1 : 0;
    rc |= std::fclose(f);
    return rc ? TDVMC_ERR_BAD_ARG : TDVMC_OK;
}

}  // extern "C"
