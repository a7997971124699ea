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

#include <cstdio>
#include <exception>
#include <string>

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

}  // namespace

extern "C" {

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
