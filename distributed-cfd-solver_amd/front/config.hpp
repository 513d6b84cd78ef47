// The solver's two configuration files, read as the reference reads them.
//
// parse_solver_config mirrors src/solver/compressible/config.cpp:28-288 and
// parse_boundary_config src/bc/config.cpp:21-289: the same sections, keys,
// defaults, conversions (toml++ value rules, front/toml.hpp), validation
// order and messages.  Where the reference calls mpi::fatal(comm, what) and
// aborts every rank, these functions throw ConfigError(what) -- the message
// text is the reference's, "config: ..." included.  The structures carry
// the reference's enumerations in the reference's order
// (include/cfd/solver/compressible/config.hpp, include/cfd/linalg/config.hpp,
// include/cfd/bc/config.hpp, physical_bc.hpp).
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace cfdg::front {

struct ConfigError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// constants of include/cfd/core/types.hpp:36-44
constexpr double kIsaPressure = 101325.0, kIsaTemperature = 288.15, kIsaDensity = 1.225;
constexpr double kAirGamma = 1.4, kAirGasConstant = 287.052874, kAirPrandtl = 0.71;

enum class EqOfState : int { IdealGas = 0 };
enum class FlowModel : int { Inviscid = 0, Viscous = 1 };
enum class Flux : int { HLLC = 0 };
enum class TimeScheme : int { ForwardEuler = 0, BackwardEuler = 1 };
enum class TimeMode : int { Steady = 0, Unsteady = 1 };
enum class Recon : int { FirstOrder = 0, Muscl = 1, MusclDirectional = 2 };
enum class Limiter : int { None = 0, Venkatakrishnan = 1, BarthJespersen = 2, VanAlbada = 3, Minmod1D = 4,
                           VanAlbada1D = 5 };
enum class Gradient : int { GreenGaussFace = 0, GreenGaussCell = 1, LeastSquaresCellFace = 2,
                            LeastSquaresCellNode = 3 };
enum class LinSolver : int { BiCGSTAB = 0 };
enum class Precond : int { None = 0, SGS = 1 };
enum class Verbosity : int { Silent = 0, Summary = 1, Verbose = 2 };

struct SolverConfig {
    // [flow]
    EqOfState eos = EqOfState::IdealGas;
    double gamma = kAirGamma, gas_constant = kAirGasConstant;
    double prandtl = kAirPrandtl;
    FlowModel flow_model = FlowModel::Inviscid;
    // [initial]
    double init_rho = kIsaDensity, init_p = kIsaPressure;
    std::array<double, 3> init_velocity{0.0, 0.0, 0.0};
    // [numerics]
    Gradient gradient = Gradient::GreenGaussFace;
    Flux flux = Flux::HLLC;
    Recon reconstruction = Recon::FirstOrder;
    Limiter limiter = Limiter::Venkatakrishnan;
    double venkat_k = 0.5;
    // [turbulence]
    bool turbulence = false;  // model = "SA"
    double nu_inf_ratio = 3.0;
    int max_distance_sweeps = 500;
    double distance_tolerance = 1.0e-8;
    // [time]
    TimeScheme scheme = TimeScheme::ForwardEuler;
    double cfl = 0.4;
    std::int64_t max_iterations = 10000;
    double residual_tolerance = 1.0e-10;
    TimeMode time_mode = TimeMode::Steady;
    double dt = 1e-6;
    std::int64_t max_time_steps = 1;
    std::int64_t bdf_order = 2;
    // [linalg] (SolverParams defaults when the section is absent)
    LinSolver solver = LinSolver::BiCGSTAB;
    Precond preconditioner = Precond::None;
    double rel_tol = 1.0e-1, abs_tol = 1.0e-15;
    int max_iter = 500;
    Verbosity verbosity = Verbosity::Silent;
    // [output]
    std::string output_dir = "out/solver";
    std::int64_t field_interval = 200;
    std::int64_t residual_interval = 20;
};

// BCType / InflowMode of include/cfd/bc/physical/physical_bc.hpp (the same
// numbering as CFDG_BC_* / CFDG_INFLOW_* of cfdg.h)
enum class BcType : int { SupersonicInlet = 0, SupersonicOutlet = 1, SubsonicInlet = 2, SubsonicOutlet = 3,
                          SlipWall = 4, NoSlipWall = 5, NoSlipWallHeatFlux = 6, Symmetry = 7, Farfield = 8 };
enum class InflowMode : int { Velocity = 0, MachAngles = 1, MachDirection = 2 };

struct BcDescriptor {
    BcType type = BcType::Symmetry;
    InflowMode inflow_mode = InflowMode::Velocity;
    int patch_id = -1;
    double p = kIsaPressure, t = kIsaTemperature, tmp_grad = 0.0;
    std::array<double, 3> velocity{0.0, 0.0, 0.0};
    double mach = 0.0, alpha_deg = 0.0, beta_deg = 0.0;
    std::array<double, 3> direction{1.0, 0.0, 0.0};
};

struct BoundaryConfig {
    std::vector<BcDescriptor> patches;  // one per mesh patch, indexed by patch id
    std::vector<std::string> warnings;  // the reference's log_stat warnings (verbose runs print them)
};

// The file is read whole; `text` overloads parse a document given in memory
// with `path` used in the messages.
SolverConfig parse_solver_config(const std::string& path);
SolverConfig parse_solver_config_text(const std::string& text, const std::string& path);
// patch_names: the mesh's patches in id order (MeshPart::patches[p].name)
BoundaryConfig parse_boundary_config(const std::string& path, const std::vector<std::string>& patch_names);
BoundaryConfig parse_boundary_config_text(const std::string& text, const std::string& path,
                                          const std::vector<std::string>& patch_names);

}  // namespace cfdg::front
