#include "config.hpp"

#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace tdvmc {

namespace {

enum class ItemType { String, Int, Int64, Double, ArrInt, ArrDouble };

struct Item {
    const char* name;
    ItemType type;
    void* ptr;
};

// ConfigItem::setValue (src/ConfigItem.cpp:45-80): the value is applied unless
// it compares equal to the integer 0 -- so an explicit integer 0 keeps the
// previous value while a missing key (null) resets to 0 / empty.
void apply(const Item& it, const JsonValue* v)
{
    static const JsonValue null_value;
    const JsonValue& val = v ? *v : null_value;
    if (val.type == JsonValue::Int && val.i == 0) return;
    switch (it.type) {
    case ItemType::String:
        *static_cast<std::string*>(it.ptr) = val.type == JsonValue::String ? val.s : std::string();
        break;
    case ItemType::Int:
        *static_cast<int*>(it.ptr) = (int)val.as_int();
        break;
    case ItemType::Int64:
        *static_cast<int64_t*>(it.ptr) = val.as_int();
        break;
    case ItemType::Double:
        *static_cast<double*>(it.ptr) = val.as_double();
        break;
    case ItemType::ArrInt: {
        auto* out = static_cast<std::vector<int>*>(it.ptr);
        out->clear();
        for (const auto& e : val.arr) out->push_back((int)e.as_int());
        break;
    }
    case ItemType::ArrDouble: {
        auto* out = static_cast<std::vector<double>*>(it.ptr);
        out->clear();
        for (const auto& e : val.arr) out->push_back(e.as_double());
        break;
    }
    }
}

// The config items in the reference's order (src/TDVMC.cpp:297-349), then
// this library's keys
std::vector<Item> items_of(Config& c)
{
    return {
        {"CONFIG_VERSION", ItemType::String, &c.CONFIG_VERSION},
        {"SYSTEM_TYPE", ItemType::String, &c.SYSTEM_TYPE},
        {"OUTPUT_DIRECTORY", ItemType::String, &c.OUTPUT_DIRECTORY},
        {"OUT_DIR_SUFFIX", ItemType::String, &c.OUT_DIR_SUFFIX},
        {"OUT_DIR_NAME", ItemType::String, &c.OUT_DIR_NAME},
        {"N", ItemType::Int, &c.N},
        {"LBOX", ItemType::Double, &c.LBOX},
        {"DIM", ItemType::Int, &c.DIM},
        {"N_PARAM", ItemType::Int, &c.N_PARAM},
        {"USE_PARAM_START", ItemType::Int, &c.USE_PARAM_START},
        {"USE_PARAM_END", ItemType::Int, &c.USE_PARAM_END},
        {"RHO", ItemType::Double, &c.RHO},
        {"RC", ItemType::Double, &c.RC},
        {"MC_STEP", ItemType::Double, &c.MC_STEP},
        {"MC_STEP_OFFSET", ItemType::Double, &c.MC_STEP_OFFSET},
        {"MC_NSTEPS", ItemType::Int, &c.MC_NSTEPS},
        {"MC_NTHERMSTEPS", ItemType::Int, &c.MC_NTHERMSTEPS},
        {"MC_NINITIALIZATIONSTEPS", ItemType::Int, &c.MC_NINITIALIZATIONSTEPS},
        {"MC_VERY_FIRST_NINITIALIZATIONSTEPS", ItemType::Int, &c.MC_VERY_FIRST_NINITIALIZATIONSTEPS},
        {"MC_NADDITIONALSTEPS", ItemType::Int, &c.MC_NADDITIONALSTEPS},
        {"MC_NADDITIONALTHERMSTEPS", ItemType::Int, &c.MC_NADDITIONALTHERMSTEPS},
        {"MC_NADDITIONALINITIALIZATIONSTEPS", ItemType::Int, &c.MC_NADDITIONALINITIALIZATIONSTEPS},
        {"MC_NFINALSTEPS_MULTIPLICATOR", ItemType::Int, &c.MC_NFINALSTEPS_MULTIPLICATOR},
        {"TIMESTEP", ItemType::Double, &c.TIMESTEP},
        {"TOTALTIME", ItemType::Double, &c.TOTALTIME},
        {"IMAGINARY_TIME", ItemType::Int, &c.IMAGINARY_TIME},
        {"ODE_SOLVER_TYPE", ItemType::Int, &c.ODE_SOLVER_TYPE},
        {"LINEAR_EQUATION_SOLVER_TYPE", ItemType::Int, &c.LINEAR_EQUATION_SOLVER_TYPE},
        {"USE_PRECONDITIONING", ItemType::Int, &c.USE_PRECONDITIONING},
        {"USE_PARAMETER_ACCEPTANCE_CHECK", ItemType::Int, &c.USE_PARAMETER_ACCEPTANCE_CHECK},
        {"PARAMETER_ACCEPTANCE_CHECK_TYPE", ItemType::Int, &c.PARAMETER_ACCEPTANCE_CHECK_TYPE},
        {"WRITE_EVERY_NTH_STEP_TO_FILE", ItemType::Int, &c.WRITE_EVERY_NTH_STEP_TO_FILE},
        {"CALCULATE_ADDITIONAL_DATA_EVERY_NTH_STEP", ItemType::Int, &c.CALCULATE_ADDITIONAL_DATA_EVERY_NTH_STEP},
        {"WRITE_SINGLE_FILES", ItemType::Int, &c.WRITE_SINGLE_FILES},
        {"MC_NSTEP_MULTIPLICATION_FACTOR_FOR_WRITE_DATA", ItemType::Int, &c.MC_NSTEP_MULTIPLICATION_FACTOR_FOR_WRITE_DATA},
        {"USE_MEAN_FOR_FINAL_PARAMETERS", ItemType::Int, &c.USE_MEAN_FOR_FINAL_PARAMETERS},
        {"USE_NORMALIZE_WF", ItemType::Int, &c.USE_NORMALIZE_WF},
        {"USE_ADJUST_PARAMETERS", ItemType::Int, &c.USE_ADJUST_PARAMETERS},
        {"UPDATE_SAMPLES_EVERY_NTH_STEP", ItemType::Int, &c.UPDATE_SAMPLES_EVERY_NTH_STEP},
        {"UPDATE_SAMPLES_PERCENT", ItemType::Double, &c.UPDATE_SAMPLES_PERCENT},
        {"GR_BIN_COUNT", ItemType::Int, &c.GR_BIN_COUNT},
        {"RHO_BIN_COUNT", ItemType::Int, &c.RHO_BIN_COUNT},
        {"USE_NURBS", ItemType::Int, &c.USE_NURBS},
        {"NURBS_GRID", ItemType::ArrDouble, &c.NURBS_GRID},
        {"PARTICLE_TYPES", ItemType::ArrInt, &c.PARTICLE_TYPES},
        {"SYSTEM_PARAMS", ItemType::ArrDouble, &c.SYSTEM_PARAMS},
        {"PARAMS_REAL", ItemType::ArrDouble, &c.PARAMS_REAL},
        {"PARAMS_IMAGINARY", ItemType::ArrDouble, &c.PARAMS_IMAGINARY},
        {"PARAM_PHIR", ItemType::Double, &c.PARAM_PHIR},
        {"PARAM_PHII", ItemType::Double, &c.PARAM_PHII},
        // library keys
        {"HBAR2_2M", ItemType::Double, &c.HBAR2_2M},
        {"PAIR_POTENTIAL", ItemType::String, &c.PAIR_POTENTIAL},
        {"K_VECTORS", ItemType::Int, &c.K_VECTORS},
        {"ONE_BODY", ItemType::Int, &c.ONE_BODY},
        {"TRAP_OMEGA", ItemType::Double, &c.TRAP_OMEGA},
        {"SPLINE_CUTOFF", ItemType::Double, &c.SPLINE_CUTOFF},
        {"N_WALKERS", ItemType::Int, &c.N_WALKERS},
        {"SAMPLES_PER_WALKER", ItemType::Int, &c.SAMPLES_PER_WALKER},
        {"SEED", ItemType::Int64, &c.SEED},
        {"PRECISION", ItemType::String, &c.PRECISION},
        {"REG_EPS", ItemType::Double, &c.REG_EPS},
        {"MOVE_ORDER", ItemType::String, &c.MOVE_ORDER},
        {"SPLINE_BC_ORIGIN", ItemType::Int, &c.SPLINE_BC_ORIGIN},
        {"SPLINE_BC_CUTOFF", ItemType::Int, &c.SPLINE_BC_CUTOFF},
    };
}

// ConfigItem::getJsonString (src/ConfigItem.cpp:82-106): strings quoted,
// integers and doubles by std::to_string (doubles therefore with six decimals),
// arrays on their own lines with JoinVector's comma-joined entries (doubles in
// scientific notation with 15 digits, src/Utils.h:56-72)
std::string json_value(const Item& it)
{
    char buf[64];
    std::string s;
    switch (it.type) {
    case ItemType::String:
        return "\"" + *static_cast<const std::string*>(it.ptr) + "\"";
    case ItemType::Int:
        return std::to_string(*static_cast<const int*>(it.ptr));
    case ItemType::Int64:
        return std::to_string(*static_cast<const int64_t*>(it.ptr));
    case ItemType::Double:
        return std::to_string(*static_cast<const double*>(it.ptr));
    case ItemType::ArrInt:
        for (int v : *static_cast<const std::vector<int>*>(it.ptr)) s += (s.empty() ? "" : ",") + std::to_string(v);
        break;
    case ItemType::ArrDouble:
        for (double v : *static_cast<const std::vector<double>*>(it.ptr)) {
            std::snprintf(buf, sizeof(buf), "%.15e", v);
            s += (s.empty() ? "" : ",") + std::string(buf);
        }
        break;
    }
    return "\n\t[\n\t\t " + s + "\n\t]";
}

}  // namespace

std::string config_to_json(const Config& cfg)
{
    Config c = cfg;
    const std::vector<Item> items = items_of(c);
    // WriteConfig (src/TDVMC.cpp:387-410)
    std::string out = "{\n";
    for (size_t i = 0; i < items.size(); i++)
        out += "\t\"" + std::string(items[i].name) + "\" : " + json_value(items[i]) + (i + 1 < items.size() ? ",\n" : "\n");
    return out + "}";
}

namespace {
}  // namespace

Config parse_config(const std::string& text)
{
    JsonValue root = JsonParser(text).parse();
    if (root.type != JsonValue::Object) throw std::runtime_error("config is not a JSON object");
    Config c;
    const std::vector<Item> items = items_of(c);
    for (const auto& it : items) apply(it, root.get(it.name));
    return c;
}

std::string read_config_text(const std::string& arg)
{
    size_t p = arg.find_first_not_of(" \t\r\n");
    if (p != std::string::npos && arg[p] == '{') return arg;
    std::ifstream f(arg, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open config file '" + arg + "'");
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

}  // namespace tdvmc
