// Config schema of the reference (src/TDVMC.cpp:297-349, RegisterAllConfigItems)
// plus the keys this library adds for the k-space term, the one-body factor,
// the trap, the pair-potential choice and the batched walkers.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "json.hpp"

namespace tdvmc {

struct Config {
    // ---- reference keys (names, types and order of src/TDVMC.cpp:297-349)
    std::string CONFIG_VERSION, SYSTEM_TYPE, OUTPUT_DIRECTORY, OUT_DIR_SUFFIX, OUT_DIR_NAME;
    int N = 0;
    double LBOX = 0;
    int DIM = 0, N_PARAM = 0, USE_PARAM_START = 0, USE_PARAM_END = 0;
    double RHO = 0, RC = 0, MC_STEP = 0, MC_STEP_OFFSET = 0;
    int MC_NSTEPS = 0, MC_NTHERMSTEPS = 0, MC_NINITIALIZATIONSTEPS = 0, MC_VERY_FIRST_NINITIALIZATIONSTEPS = 0;
    int MC_NADDITIONALSTEPS = 0, MC_NADDITIONALTHERMSTEPS = 0, MC_NADDITIONALINITIALIZATIONSTEPS = 0;
    int MC_NFINALSTEPS_MULTIPLICATOR = 0;
    double TIMESTEP = 0, TOTALTIME = 0;
    int IMAGINARY_TIME = 0, ODE_SOLVER_TYPE = 0, LINEAR_EQUATION_SOLVER_TYPE = 0, USE_PRECONDITIONING = 0;
    int USE_PARAMETER_ACCEPTANCE_CHECK = 0, PARAMETER_ACCEPTANCE_CHECK_TYPE = 0, WRITE_EVERY_NTH_STEP_TO_FILE = 0;
    int CALCULATE_ADDITIONAL_DATA_EVERY_NTH_STEP = 0, WRITE_SINGLE_FILES = 0;
    int MC_NSTEP_MULTIPLICATION_FACTOR_FOR_WRITE_DATA = 0, USE_MEAN_FOR_FINAL_PARAMETERS = 0;
    int USE_NORMALIZE_WF = 0, USE_ADJUST_PARAMETERS = 0, UPDATE_SAMPLES_EVERY_NTH_STEP = 0;
    double UPDATE_SAMPLES_PERCENT = 0;
    int GR_BIN_COUNT = 0, RHO_BIN_COUNT = 0, USE_NURBS = 0;
    std::vector<double> NURBS_GRID;
    std::vector<int> PARTICLE_TYPES;
    std::vector<double> SYSTEM_PARAMS, PARAMS_REAL, PARAMS_IMAGINARY;
    double PARAM_PHIR = 0, PARAM_PHII = 0;

    // ---- keys added by this library (absent from the reference)
    double HBAR2_2M = 0;            // 0 -> 1.0 (src/Constants.h:12)
    std::string PAIR_POTENTIAL;     // SQUARE_WELL | GAUSSIAN | LJ | NONE ("" -> by SYSTEM_TYPE)
    int K_VECTORS = 0;              // k-space Jastrow terms
    int ONE_BODY = 0;               // 1: one-body Gaussian factor (parameter 0)
    double TRAP_OMEGA = 0;          // external 0.5 omega^2 r^2
    double SPLINE_CUTOFF = 0;       // 0 -> LBOX / 2
    int N_WALKERS = 0;              // 0 -> 1
    int SAMPLES_PER_WALKER = 0;     // 0 -> MC_NSTEPS, or 1 without it
    int64_t SEED = 0;               // 0 -> 1
    std::string PRECISION;          // FP64 | MIXED ("" -> FP64)
    double REG_EPS = 0;             // 0 -> 0.001 (src/TDVMC.cpp:1737)
    std::string MOVE_ORDER;         // RANDOM | SEQUENTIAL ("" -> RANDOM)
    int SPLINE_BC_ORIGIN = 0;       // pair-spline boundary condition at r = 0: 1..4 (0 -> 2)
    int SPLINE_BC_CUTOFF = 0;       // ... at r = rc: 1..4 (0 -> 2)
};

// Parses a config JSON text with the reference's ConfigItem::setValue
// semantics (src/ConfigItem.cpp:45-80).  Throws std::runtime_error.
Config parse_config(const std::string& text);

// The config as WriteConfig writes it (src/TDVMC.cpp:387-410): every item
// in the reference's order, then this library's keys, which the reference's
// reader ignores.
std::string config_to_json(const Config& c);

// Reads `arg` as JSON text if it starts with '{', else as a file path.
std::string read_config_text(const std::string& arg);

}  // namespace tdvmc
