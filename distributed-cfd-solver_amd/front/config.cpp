// Solver and boundary-condition configuration files (config.hpp).
//
// The checks run in the reference's order, so that a file with several
// faults reports the same first fault: src/solver/compressible/config.cpp:28-288
// and src/bc/config.cpp:21-289, with the helpers of
// src/io/toml/toml_utilities.cpp:18-220.
#include "config.hpp"

#include <fstream>
#include <initializer_list>
#include <sstream>

#include "toml.hpp"

namespace cfdg::front {

namespace {

using toml::Node;

[[noreturn]] void fail(const std::string& what) { throw ConfigError("config: " + what); }

std::string read_file(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) fail("cannot open file '" + path + "' on Rank 0");
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

toml::NodePtr parse_document(const std::string& text, const std::string& path) {
    try {
        return toml::parse(text);
    } catch (const toml::ParseError& e) {
        fail("cannot parse '" + path + "': " + e.what());
    }
}

// toml_utilities.cpp helpers ---------------------------------------------------

const Node& req_table(const Node& t, const char* key, const std::string& ctx) {
    const Node* sub = t.get(key);
    if (sub == nullptr || !sub->is_table()) fail(ctx + ": missing required section [" + key + "]");
    return *sub;
}

double req_number(const Node& t, const char* key, const std::string& ctx) {
    const Node* n = t.get(key);
    if (n == nullptr) fail(ctx + ": missing required key '" + key + "'");
    const auto v = n->as_double();
    if (!v) fail(ctx + ": key '" + key + "' must be a number");
    return *v;
}

double opt_number(const Node& t, const char* key, double def) {
    const Node* n = t.get(key);
    if (n == nullptr) return def;
    return n->as_double().value_or(def);
}

std::int64_t req_integer(const Node& t, const char* key, const std::string& ctx) {
    const Node* n = t.get(key);
    if (n == nullptr) fail(ctx + ": missing required key '" + key + "'");
    const auto v = n->as_integer();
    if (!v) fail(ctx + ": key '" + key + "' must be an integer");
    return *v;
}

std::int64_t opt_integer(const Node& t, const char* key, std::int64_t def) {
    const Node* n = t.get(key);
    if (n == nullptr) return def;
    return n->as_integer().value_or(def);
}

std::string req_string(const Node& t, const char* key, const std::string& ctx) {
    const Node* n = t.get(key);
    if (n == nullptr) fail(ctx + ": missing required key '" + key + "'");
    const auto v = n->as_string();
    if (!v) fail(ctx + ": key '" + key + "' must be a string");
    return *v;
}

std::string opt_string(const Node& t, const char* key, const std::string& def) {
    const Node* n = t.get(key);
    if (n == nullptr) return def;
    return n->as_string().value_or(def);
}

std::array<double, 3> req_vec3(const Node& t, const char* key, const std::string& ctx) {
    const Node* n = t.get(key);
    if (n == nullptr) fail(ctx + ": missing required key '" + key + "'");
    const std::string bad = ctx + ": key '" + key + "' must be an array of 3 numbers";
    if (!n->is_array() || n->array.size() != 3) fail(bad);
    std::array<double, 3> out{};
    for (std::size_t k = 0; k < 3; ++k) {
        const auto v = n->array[k]->as_double();
        if (!v) fail(bad);
        out[k] = *v;
    }
    return out;
}

// the first key (in the table's sorted order) that is not allowed
void check_allowed_keys(const Node& t, const std::vector<std::string>& allowed, const std::string& ctx) {
    for (const auto& kv : t.table) {
        bool known = false;
        for (const std::string& a : allowed)
            if (kv.first == a) {
                known = true;
                break;
            }
        if (!known) fail(ctx + ": unknown key '" + kv.first + "'");
    }
}

void check_positive(double v, const char* what, const std::string& ctx) {
    if (!(v > 0.0)) fail(ctx + ": '" + what + "' must be positive (got " + std::to_string(v) + ")");
}

// bc/config.cpp:25-137
void parse_inflow(const Node& t, BcDescriptor& d, const std::string& ctx) {
    const bool has_vel = t.contains("velocity") || t.contains("velocity_inf");
    const bool has_mach = t.contains("mach") || t.contains("mach_inf");
    const bool has_dir = t.contains("direction");
    const bool has_angles =
        t.contains("alpha") || t.contains("alpha_deg") || t.contains("beta") || t.contains("beta_deg");
    if (has_vel && (has_mach || has_dir || has_angles))
        fail(ctx + ": conflicting parameters: cannot specify direct 'velocity' together with 'mach', 'direction', "
                   "or 'alpha/beta' angles");
    if (has_dir && has_angles)
        fail(ctx + ": conflicting parameters: cannot specify both 'direction' vector and 'alpha/beta' angles");
    if (has_dir && !has_mach) fail(ctx + ": 'direction' vector requires 'mach' to be specified");
    if (has_angles && !has_mach) fail(ctx + ": 'alpha/beta' angles require 'mach' to be specified");
    if (!has_vel && !has_mach)
        fail(ctx + ": missing velocity definition: specify either 'velocity' vector or 'mach' with "
                   "angles/direction");

    const std::vector<std::string> meta = {"patch_id", "type", "name", "cgns_type", "global_face_count",
                                           "p", "p_inf", "t", "t_inf"};
    std::vector<std::string> allowed = meta;
    if (has_vel) {
        d.inflow_mode = InflowMode::Velocity;
        allowed.insert(allowed.end(), {"velocity", "velocity_inf"});
    } else if (has_dir) {
        d.inflow_mode = InflowMode::MachDirection;
        allowed.insert(allowed.end(), {"mach", "", "direction"});
    } else {
        d.inflow_mode = InflowMode::MachAngles;
        allowed.insert(allowed.end(), {"mach", "alpha", "alpha_deg", "beta", "beta_deg"});
    }
    check_allowed_keys(t, allowed, ctx);

    if (t.contains("p"))
        d.p = req_number(t, "p", ctx);
    else if (t.contains("p_inf"))
        d.p = req_number(t, "p_inf", ctx);
    else
        fail(ctx + ": missing required pressure ('p' or 'p_inf')");
    if (t.contains("t"))
        d.t = req_number(t, "t", ctx);
    else if (t.contains("t_inf"))
        d.t = req_number(t, "t_inf", ctx);
    else
        fail(ctx + ": missing required temperature ('t' or 't_inf')");
    check_positive(d.p, "p", ctx);
    check_positive(d.t, "t", ctx);

    switch (d.inflow_mode) {
        case InflowMode::Velocity:
            d.velocity = req_vec3(t, t.contains("velocity") ? "velocity" : "velocity_inf", ctx);
            break;
        case InflowMode::MachDirection: {
            d.mach = req_number(t, "mach", ctx);
            check_positive(d.mach, "mach", ctx);
            d.direction = req_vec3(t, "direction", ctx);
            const double m2 =
                d.direction[0] * d.direction[0] + d.direction[1] * d.direction[1] + d.direction[2] * d.direction[2];
            if (m2 < 1.0e-14) fail(ctx + ": 'direction' vector cannot be zero");
            break;
        }
        case InflowMode::MachAngles:
            d.mach = req_number(t, "mach", ctx);
            check_positive(d.mach, "mach", ctx);
            d.alpha_deg = t.contains("alpha_deg") ? req_number(t, "alpha_deg", ctx)
                          : t.contains("alpha")   ? req_number(t, "alpha", ctx)
                                                  : 0.0;
            d.beta_deg = t.contains("beta_deg") ? req_number(t, "beta_deg", ctx)
                         : t.contains("beta")   ? req_number(t, "beta", ctx)
                                                : 0.0;
            break;
    }
}

std::string mach_warning(const std::string& ctx) {
    return "WARNING: " + ctx +
           ": Mach-based inflow mode selected. Note: velocity magnitude is evaluated assuming Ideal Gas EOS "
           "kinematics.";
}

}  // namespace

SolverConfig parse_solver_config_text(const std::string& text, const std::string& path) {
    SolverConfig cfg;
    const toml::NodePtr doc = parse_document(text, path);
    const Node& root = *doc;
    check_allowed_keys(root, {"flow", "initial", "numerics", "time", "output", "turbulence", "linalg"},
                       "'" + path + "'");

    {  // [flow]
        const Node& t = req_table(root, "flow", path);
        const std::string ctx = path + " [flow]";
        const std::string eos = req_string(t, "eos", ctx);
        const std::string model = req_string(t, "flow_model", ctx);
        std::vector<std::string> allowed = {"eos", "flow_model"};
        if (eos == "IDEAL_GAS") {
            cfg.eos = EqOfState::IdealGas;
            allowed.insert(allowed.end(), {"gamma", "gas_constant"});
            cfg.gamma = req_number(t, "gamma", ctx);
            cfg.gas_constant = req_number(t, "gas_constant", ctx);
            check_positive(cfg.gamma, "gamma", ctx);
            check_positive(cfg.gas_constant, "gas_constant", ctx);
        } else {
            fail(ctx + ": unsupported EOS model '" + eos + "' (available: IDEAL_GAS)");
        }
        if (model == "INVISCID_FLOW") {
            cfg.flow_model = FlowModel::Inviscid;
        } else if (model == "VISCOUS_FLOW") {
            cfg.flow_model = FlowModel::Viscous;
            allowed.push_back("prandtl");
            cfg.prandtl = req_number(t, "prandtl", ctx);
            check_positive(cfg.prandtl, "prandtl", ctx);
        } else {
            fail(ctx + ": unsupported Flow Model '" + model + "' (available: INVISCID_FLOW, VISCOUS_FLOW)");
        }
        check_allowed_keys(t, allowed, ctx);
    }

    {  // [initial]
        const Node& t = req_table(root, "initial", path);
        const std::string ctx = path + " [initial]";
        check_allowed_keys(t, {"rho", "pressure", "velocity"}, ctx);
        cfg.init_rho = req_number(t, "rho", ctx);
        cfg.init_p = req_number(t, "pressure", ctx);
        cfg.init_velocity = req_vec3(t, "velocity", ctx);
        check_positive(cfg.init_rho, "rho", ctx);
        check_positive(cfg.init_p, "pressure", ctx);
    }

    {  // [numerics]
        const Node& t = req_table(root, "numerics", path);
        const std::string ctx = path + " [numerics]";
        check_allowed_keys(t, {"flux", "reconstruction", "limiter", "venkat_k", "gradient"}, ctx);
        const std::string flux = req_string(t, "flux", ctx);
        if (flux == "HLLC")
            cfg.flux = Flux::HLLC;
        else
            fail(ctx + ": unsupported flux '" + flux + "' (available: HLLC)");
        const std::string reco = req_string(t, "reconstruction", ctx);
        if (reco == "FIRST_ORDER") {
            cfg.reconstruction = Recon::FirstOrder;
            cfg.limiter = Limiter::None;
        } else if (reco == "MUSCL") {
            cfg.reconstruction = Recon::Muscl;
            const std::string lim = req_string(t, "limiter", ctx);
            if (lim == "VENKAT")
                cfg.limiter = Limiter::Venkatakrishnan;
            else if (lim == "BARTH")
                cfg.limiter = Limiter::BarthJespersen;
            else if (lim == "VAN_ALBADA")
                cfg.limiter = Limiter::VanAlbada;
            else
                fail(ctx + ": invalid limiter '" + lim + "' for MUSCL (available: VENKAT, BARTH, VAN_ALBADA)");
        } else if (reco == "MUSCL_DIRECTIONAL") {
            cfg.reconstruction = Recon::MusclDirectional;
            const std::string lim = req_string(t, "limiter", ctx);
            if (lim == "MINMOD_1D")
                cfg.limiter = Limiter::Minmod1D;
            else if (lim == "VAN_ALBADA_1D")
                cfg.limiter = Limiter::VanAlbada1D;
            else
                fail(ctx + ": invalid limiter '" + lim +
                     "' for MUSCL_DIRECTIONAL (available: MINMOD_1D, VAN_ALBADA_1D)");
        } else {
            fail(ctx + ": unsupported reconstruction '" + reco +
                 "' (available: FIRST_ORDER, MUSCL, MUSCL_DIRECTIONAL)");
        }
        cfg.venkat_k = opt_number(t, "venkat_k", 0.5);
        check_positive(cfg.venkat_k, "venkat_k", ctx);
        const std::string g = opt_string(t, "gradient", "GREEN_GAUSS_FACE");
        if (g == "GREEN_GAUSS_FACE")
            cfg.gradient = Gradient::GreenGaussFace;
        else if (g == "GREEN_GAUSS_CELL")
            cfg.gradient = Gradient::GreenGaussCell;
        else if (g == "LEAST_SQUARES_FACE" || g == "LEAST_SQUARES_NODE")  // both: the node operator (config.cpp:144-147)
            cfg.gradient = Gradient::LeastSquaresCellNode;
        else
            fail(ctx + ": unsupported gradient '" + g +
                 "' (available: GREEN_GAUSS_FACE, GREEN_GAUSS_CELL, LEAST_SQUARES_FACE, LEAST_SQUARES_NODE)");
    }

    if (const Node* n = root.get("turbulence")) {  // [turbulence], optional
        if (!n->is_table()) fail(path + ": [turbulence] must be a table");
        const Node& t = *n;
        const std::string ctx = path + " [turbulence]";
        check_allowed_keys(t, {"model", "nu_inf_ratio", "max_distance_sweeps", "distance_tolerance"}, ctx);
        const std::string model = req_string(t, "model", ctx);
        if (model == "SA")
            cfg.turbulence = true;
        else
            fail(ctx + ": unsupported turbulence model '" + model + "' (available: SA)");
        cfg.nu_inf_ratio = opt_number(t, "nu_inf_ratio", 3.0);
        check_positive(cfg.nu_inf_ratio, "nu_inf_ratio", ctx);
        cfg.max_distance_sweeps = static_cast<int>(opt_integer(t, "max_distance_sweeps", 500));
        cfg.distance_tolerance = opt_number(t, "distance_tolerance", 1.0e-8);
        check_positive(cfg.distance_tolerance, "distance_tolerance", ctx);
        if (cfg.turbulence && cfg.flow_model != FlowModel::Viscous)
            fail(ctx + ": turbulence requires [numerics] viscous = true");
    }

    {  // [time]
        const Node& t = req_table(root, "time", path);
        const std::string ctx = path + " [time]";
        check_allowed_keys(t, {"scheme", "cfl", "max_iterations", "residual_tolerance", "mode", "time_step",
                               "max_time_steps", "bdf_order"},
                           ctx);
        const std::string scheme = req_string(t, "scheme", ctx);
        if (scheme == "FORWARD_EULER")
            cfg.scheme = TimeScheme::ForwardEuler;
        else if (scheme == "BACKWARD_EULER")
            cfg.scheme = TimeScheme::BackwardEuler;
        else
            fail(ctx + ": unsupported scheme '" + scheme + "' (available: FORWARD_EULER, BACKWARD_EULER)");
        cfg.cfl = req_number(t, "cfl", ctx);
        check_positive(cfg.cfl, "cfl", ctx);
        cfg.max_iterations = req_integer(t, "max_iterations", ctx);
        if (cfg.max_iterations < 1) fail(ctx + ": 'max_iterations' must be >= 1");
        cfg.residual_tolerance = req_number(t, "residual_tolerance", ctx);
        check_positive(cfg.residual_tolerance, "residual_tolerance", ctx);
        const std::string mode = req_string(t, "mode", ctx);
        if (mode == "Steady") {
            cfg.time_mode = TimeMode::Steady;
        } else if (mode == "Unsteady") {
            cfg.time_mode = TimeMode::Unsteady;
            cfg.dt = req_number(t, "time_step", ctx);
            cfg.max_time_steps = req_integer(t, "max_time_steps", ctx);
            cfg.bdf_order = opt_integer(t, "bdf_order", 2);
        }  // any other mode keeps the steady default, as the reference does
    }

    {  // [output]
        const Node& t = req_table(root, "output", path);
        const std::string ctx = path + " [output]";
        check_allowed_keys(t, {"directory", "field_interval", "residual_interval"}, ctx);
        cfg.output_dir = req_string(t, "directory", ctx);
        cfg.field_interval = req_integer(t, "field_interval", ctx);
        cfg.residual_interval = req_integer(t, "residual_interval", ctx);
        if (cfg.field_interval < 0 || cfg.residual_interval < 1)
            fail(ctx + ": intervals must be >= 0 (field) / >= 1 (residual)");
    }

    if (const Node* n = root.get("linalg")) {  // [linalg], optional
        if (!n->is_table()) fail(path + ": [linalg] must be a table");
        const Node& t = *n;
        const std::string ctx = path + " [linalg]";
        check_allowed_keys(t, {"solver", "preconditioner", "rel_tol", "abs_tol", "max_iter", "verbosity", "res_verify"},
                           ctx);
        const std::string type = opt_string(t, "solver", "BICGSTAB");
        if (type == "BICGSTAB")
            cfg.solver = LinSolver::BiCGSTAB;
        else
            fail(ctx + ": unsupported solver type '" + type + "' (available: BICGSTAB)");
        const std::string pc = opt_string(t, "preconditioner", "None");
        if (pc == "None")
            cfg.preconditioner = Precond::None;
        else if (pc == "SGS")
            cfg.preconditioner = Precond::SGS;
        else
            fail(ctx + ": unsupported preconditioner type '" + pc + "' (available: None, SGS)");
        cfg.rel_tol = opt_number(t, "rel_tol", 1e-1);
        cfg.abs_tol = opt_number(t, "abs_tol", 1e-30);
        cfg.max_iter = static_cast<int>(opt_integer(t, "max_iter", 100));
        const std::string vb = opt_string(t, "verbosity", "Silent");
        if (vb == "Silent")
            cfg.verbosity = Verbosity::Silent;
        else if (vb == "Summary")
            cfg.verbosity = Verbosity::Summary;
        else if (vb == "Verbose")
            cfg.verbosity = Verbosity::Verbose;
        else
            fail(ctx + ": unsupported verbosity type '" + vb + "' (available: Silent, Summary, Verbose)");
        check_positive(cfg.rel_tol, "rel_tol", ctx);
        check_positive(cfg.abs_tol, "abs_tol", ctx);
        check_positive(cfg.max_iter, "max_iter", ctx);
    }
    return cfg;
}

SolverConfig parse_solver_config(const std::string& path) { return parse_solver_config_text(read_file(path), path); }

BoundaryConfig parse_boundary_config_text(const std::string& text, const std::string& path,
                                          const std::vector<std::string>& patch_names) {
    const toml::NodePtr doc = parse_document(text, path);
    const Node& root = *doc;
    check_allowed_keys(root, {"boundary_condition"}, "'" + path + "'");
    const Node* arr = root.get("boundary_condition");
    if (arr == nullptr || !arr->is_array() || arr->array.empty())
        fail("'" + path + "': no [[boundary_condition]] entries");

    const std::size_t np = patch_names.size();
    BoundaryConfig out;
    out.patches.resize(np);
    std::vector<char> seen(np, 0);
    const std::vector<std::string> meta = {"patch_id", "type", "name", "cgns_type", "global_face_count"};
    const auto with = [&](std::initializer_list<const char*> more) {
        std::vector<std::string> a = meta;
        for (const char* k : more) a.emplace_back(k);
        return a;
    };

    std::size_t idx = 0;
    for (const toml::NodePtr& item : arr->array) {
        const std::string ctx = "'" + path + "' boundary_condition[" + std::to_string(idx) + "]";
        if (!item->is_table()) fail(ctx + " must be a TOML table");
        Node& t = *item;
        const std::int64_t pid = req_integer(t, "patch_id", ctx);
        if (pid < 0 || pid >= static_cast<std::int64_t>(np))
            fail(ctx + ": patch_id " + std::to_string(pid) + " outside mesh range [0, " + std::to_string(np) + ")");
        if (seen[static_cast<std::size_t>(pid)] != 0)
            fail(ctx + ": duplicate condition for patch " + std::to_string(pid));
        seen[static_cast<std::size_t>(pid)] = 1;
        BcDescriptor& d = out.patches[static_cast<std::size_t>(pid)];
        d.patch_id = static_cast<int>(pid);
        const std::string type = req_string(t, "type", ctx);
        if (type == "SUPERSONIC_INLET") {
            d.type = BcType::SupersonicInlet;
            parse_inflow(t, d, ctx);
        } else if (type == "FARFIELD") {
            d.type = BcType::Farfield;
            parse_inflow(t, d, ctx);
            if (d.inflow_mode == InflowMode::MachAngles || d.inflow_mode == InflowMode::MachDirection)
                out.warnings.push_back(mach_warning(ctx));
        } else if (type == "SUPERSONIC_OUTLET") {
            d.type = BcType::SupersonicOutlet;
            check_allowed_keys(t, meta, ctx);
        } else if (type == "SLIP_WALL") {
            d.type = BcType::SlipWall;
            check_allowed_keys(t, meta, ctx);
        } else if (type == "SYMMETRY") {
            d.type = BcType::Symmetry;
            check_allowed_keys(t, meta, ctx);
        } else if (type == "NO_SLIP_WALL") {
            d.type = BcType::NoSlipWall;
            check_allowed_keys(t, with({"t", "t_wall", "velocity"}), ctx);
            d.t = t.contains("t_wall") ? req_number(t, "t_wall", ctx)
                  : t.contains("t")    ? req_number(t, "t", ctx)
                                       : kIsaTemperature;  // isothermal wall at the ISA temperature
            check_positive(d.t, "t", ctx);
            d.velocity = t.contains("velocity") ? req_vec3(t, "velocity", ctx) : std::array<double, 3>{0.0, 0.0, 0.0};
        } else if (type == "NO_SLIP_WALL_HEAT_FLUX" || type == "NO_SLIP_WALL_ADIABATIC") {
            d.type = BcType::NoSlipWallHeatFlux;
            check_allowed_keys(t, with({"tmp_grad", "heat_flux_grad", "velocity"}), ctx);
            d.tmp_grad = t.contains("tmp_grad")         ? req_number(t, "tmp_grad", ctx)
                         : t.contains("heat_flux_grad") ? req_number(t, "heat_flux_grad", ctx)
                                                        : 0.0;  // adiabatic
            d.velocity = t.contains("velocity") ? req_vec3(t, "velocity", ctx) : std::array<double, 3>{0.0, 0.0, 0.0};
        } else if (type == "SUBSONIC_INLET") {
            d.type = BcType::SubsonicInlet;
            if (!t.contains("p") && !t.contains("p_inf")) {
                // a placeholder pressure for the inflow parser, written into the table (bc/config.cpp:244-247)
                auto p = std::make_shared<Node>();
                p->kind = Node::Kind::Float;
                p->f = kIsaPressure;
                t.table.emplace("p", std::move(p));
            }
            parse_inflow(t, d, ctx);
            if (d.inflow_mode == InflowMode::MachAngles || d.inflow_mode == InflowMode::MachDirection)
                out.warnings.push_back(mach_warning(ctx));
        } else if (type == "SUBSONIC_OUTLET") {
            d.type = BcType::SubsonicOutlet;
            check_allowed_keys(t, with({"p", "p_outlet", "p_back"}), ctx);
            if (t.contains("p_back"))
                d.p = req_number(t, "p_back", ctx);
            else if (t.contains("p_outlet"))
                d.p = req_number(t, "p_outlet", ctx);
            else if (t.contains("p"))
                d.p = req_number(t, "p", ctx);
            else
                fail(ctx + ": missing required backpressure ('p', 'p_back', or 'p_outlet')");
            check_positive(d.p, "backpressure", ctx);
        } else {
            fail(ctx + ": unknown BC type '" + type + "'");
        }
        ++idx;
    }
    for (std::size_t p = 0; p < np; ++p)
        if (seen[p] == 0)
            fail("'" + path + "': no condition assigned to patch " + std::to_string(p) + " ('" + patch_names[p] + "')");
    return out;
}

BoundaryConfig parse_boundary_config(const std::string& path, const std::vector<std::string>& patch_names) {
    return parse_boundary_config_text(read_file(path), path, patch_names);
}

}  // namespace cfdg::front
